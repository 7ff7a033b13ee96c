// Internal kernel interfaces of the somax HIP library.
//
// A batch holds the members that share one grid: their fields are stacked as
// P planes of Ny x Nx (member m owns planes first[m] .. first[m] + nl[m] - 1),
// and every kernel runs over all planes of the batch in one launch, reading
// the member and plane tables below.  A single model is the one-member batch.
#pragma once

#include "somaxhip.h"
#include "common.h"

struct somaxhip_qg_plan_s;

namespace somax {

constexpr int MAX_LAYERS = 8;

// Workgroups are dispatched to the 8 XCDs round-robin by linear ID (x
// fastest, then y, then z).  xcd_tile3 remaps the linear ID of a 3-D grid so
// that each XCD takes one contiguous run of tiles in (x, y, z) order: tiles
// that share halo rows then share that XCD's L2.
__device__ inline void xcd_tile3(int& bx, int& by, int& bz) {
  const unsigned nx = gridDim.x, ny = gridDim.y;
  const unsigned T = nx * ny * gridDim.z;
  const unsigned b = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const unsigned xcd = b & 7u, slot = b >> 3, per = T >> 3, extra = T & 7u;
  const unsigned t = xcd * per + (xcd < extra ? xcd : extra) + slot;
  bx = (int)(t % nx);
  by = (int)((t / nx) % ny);
  bz = (int)(t / (nx * ny));
}

// Merged launches.  One launch may cover the blocks of up to MAX_GROUP
// batches (the grid classes of an ensemble), so that a kernel phase of a
// Runge-Kutta stage is one launch for all of them: batch c owns the linear
// blocks [start[c], start[c] + count[c]) and sees them as its own
// gx x gy x gz grid, x fastest.  Every start is a multiple of 8, so a batch's
// block l runs on the XCD that its own launch would give it (blocks are dealt
// to the 8 XCDs round-robin); the blocks between count[c] and the next start
// exit at once.  Each batch's arithmetic is exactly that of its own launch.
constexpr int MAX_GROUP = 8;
struct BlockTable {
  int n;
  int start[MAX_GROUP + 1];
  int count[MAX_GROUP];
  int gx[MAX_GROUP], gy[MAX_GROUP], gxy[MAX_GROUP];
  float rgx[MAX_GROUP], rgxy[MAX_GROUP];  // 1 / gx, 1 / (gx * gy)
};
inline void table_add(BlockTable& t, int gx, int gy, int gz) {
  const int count = gx * gy * gz;  // below 2^24 (exact in FP32), as every launch here
  t.gx[t.n] = gx;
  t.gy[t.n] = gy;
  t.gxy[t.n] = gx * gy;
  t.rgx[t.n] = 1.0f / (float)gx;
  t.rgxy[t.n] = 1.0f / (float)(gx * gy);
  t.count[t.n] = count;
  t.start[t.n + 1] = t.start[t.n] + (count + 7) / 8 * 8;
  ++t.n;
}
// Batch c of linear block b and the block's index l in that batch; false for
// the padding blocks between batches.  The starts are read with constant
// indices, so they arrive in one batch of scalar loads.
__device__ inline bool table_class(const BlockTable& t, int b, int& c, int& l) {
  c = 0;
#pragma unroll
  for (int i = 1; i < MAX_GROUP; ++i) c += (i < t.n && b >= t.start[i]) ? 1 : 0;
  l = b - t.start[c];
  return l < t.count[c];
}
// v / d for 0 <= v < 2^24 from the FP32 reciprocal r = 1/d (one correction step)
__device__ inline int small_div(int v, int d, float r) {
  int q = (int)((float)v * r);
  if (q * d > v) --q;
  else if ((q + 1) * d <= v) ++q;
  return q;
}
// Block (x, y, z) of batch c's grid for its block index l.
__device__ inline void class_block(const BlockTable& t, int c, int l, int& x, int& y, int& z) {
  z = small_div(l, t.gxy[c], t.rgxy[c]);
  const int r = l - z * t.gxy[c];
  y = small_div(r, t.gx[c], t.rgx[c]);
  x = r - y * t.gx[c];
}
// Block b of T dealt to the XCDs round-robin -> the index that hands each XCD
// one contiguous run (the order xcd_tile3 gives a 3-D grid).
__device__ inline int xcd_order(int b, int T) {
  const unsigned xcd = (unsigned)b & 7u, slot = (unsigned)b >> 3, per = (unsigned)T >> 3, extra = (unsigned)T & 7u;
  return (int)(xcd * per + (xcd < extra ? xcd : extra) + slot);
}

// Planes, rows and columns (ghosts included) of a QG plan's fields
// (SOMAXHIP_ERR_ARG for an FP64 plan, whose fields are doubles).
int qg_plan_shape(somaxhip_qg_plan_s* plan, int* planes, int* Ny, int* Nx);

// The FP64 form of a QG model (qg64.hip): built by a plan whose description
// asks for SOMAXHIP_PREC_FP64, used by the *_f64 entry points.
struct Qg64;
Qg64* qg64_create(const somaxhip_qg_desc* d, std::string* err);
void qg64_destroy(Qg64* p);
size_t qg64_workspace_bytes(const Qg64* p);
size_t qg64_tsit5_workspace_bytes(const Qg64* p);
// The plan's FP64 form; null for an FP32 / BF16 plan.
const Qg64* qg_plan_f64(somaxhip_qg_plan_s* plan);

// Member and plane tables of a batch (device arrays).
struct Tables {
  // per member (M entries)
  const int* mfirst;           // first plane
  const int* mnl;              // layer count
  const double* cl2m;          // [M*64] Cl2m of the member, row p at p*8 (mode p from layers)
  const float* cm2l;           // [M*64] Cm2l of the member, row l at l*8 (layer l from modes)
  const unsigned char* mask;   // 1: ghost ring of psi set to 0 after the solve
  // per plane (P entries)
  const int* member;           // plane -> member
  const int* layer;            // plane -> layer (or mode) index within the member
  const float* lam;            // Helmholtz parameter of the plane's mode
  const unsigned char* hp;     // 1: Poisson-like mode (high-precision inverse)
  const float* wind_scale;     // multiplier of tau0 * wind pattern for the layer
  const float* psi_wall;       // WENO wall value of psi (null: 0)
};

// Physical parameters of one member: nu, nu4, kappa, tau0.
struct MemberParams {
  float nu, nu4, kappa, tau0;
};

// O[a][b] = sum_k A[a][k] B[b][k] with the data plane X in the A slot
// (trans == 0: O = X S) or in the B slot (trans == 1: O = (X S)^T), S a sine
// matrix (rows = outputs, k contiguous).  FP32 operands when Xf != null, else
// bf16 pieces.
//
// Folded transforms (fold_P > 0): with S[N-1-n][a] = (-1)^a S[n][a], a DST-I
// of even length N = 2h splits into two independent half-length products by
// output parity, and a 2-D transform into four blocks b = 2p + q (p: parity
// along y, q: along x).  Batch plane z is block z / fold_P of table plane
// z % fold_P; the sine operand and the eigenvalue rows come in two parity
// variants, s_var and lam_var elements apart.
//
// Fringe: when RA (RB) exceeds a multiple of 128 by 1 or 2, the 128 x 128
// tiles stop at that multiple and fringe_wgs extra workgroups per plane
// compute the leftover rows (columns) as FP64-accumulated dot products, and
// write the zero padding of columns [RB, RBpad) that no tile covers.
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
  // per plane precision: planes with hp[z] != 0 use np_lo products and write
  // op_lo output pieces, the others np_hi / op_hi (pieces 0 = FP32 out)
  const unsigned char* hp;
  int np_lo, np_hi, op_lo, op_hi;
  int acc64;  // add the FP32 partial of every few k-tiles into an FP64 running sum
  // output O[ro][co], ro < RA, co < RB; columns [RB, RBpad) written as 0
  float* Of;
  uint16_t* Ob;
  long ldo, o_piece, o_batch;
  int RA, RB, RBpad;
  // optional spectral epilogue: O /= (lamA[ro] + lamB[co] - lam[z])
  const float* lamA;
  const float* lamB;
  const float* lam;
  // folded transforms (see above); 0 = plain
  int fold_P;
  int s_par;    // parity choosing the sine variant: 0 = block & 1 (x), 1 = block >> 1 (y)
  long s_var;   // elements between the sine operand's parity variants
  int lam_var;   // floats between the parity variants of lamA (x parity)
  int lamB_var;  // floats between the parity variants of lamB (y parity)
  int nplanes;   // batch planes of the launch (set by launch_dst_gemm)
};

// Tile cut of a launch: rows (columns) covered by full tiles, am (bm), and
// fringe workgroups per plane, fpw.  A dimension exceeding a multiple of 128
// by 1 or 2 leaves that fringe to the fringe workgroups, whose fixed operand
// rows (at most 4, FP64) are staged in 40 KiB of LDS; otherwise tiles cover
// the whole dimension (fpw = 0).
struct GemmCut {
  int am, bm, fpw;
};
__host__ __device__ inline GemmCut gemm_cut(int RA, int RB, int K) {
  auto cut = [](int R) {
    const int r = R & 127;
    return (R > 128 && (r == 1 || r == 2)) ? R - r : R;
  };
  GemmCut c{cut(RA), cut(RB), 0};
  const long k8 = (K + 7) / 8 * 8;
  if ((long)((RA - c.am) + (RB - c.bm)) * k8 * 8 > 40960) {
    c.am = RA;
    c.bm = RB;
  }
  const long dots = (long)(RA - c.am) * RB + (long)c.am * (RB - c.bm);
  c.fpw = dots > 0 ? (int)((dots + 255) / 256) : 0;
  return c;
}

void launch_dst_gemm(const GemmArgs& a, int batch, hipStream_t st);
// One launch for n (<= MAX_GROUP) transforms of the same product count and
// operand type; transform i covers batch[i] planes.
void launch_dst_gemm_group(const GemmArgs* a, const int* batch, int n, hipStream_t st);
// Raises the bf16 kernel's dynamic LDS limit; call before the first launch.
hipError_t prepare_dst_gemm();

// Layer -> mode projection with FP64 accumulation (the stretching term makes
// layer PV much larger than the barotropic modal PV, so an FP32 sum cancels):
//   qm[first + p] = sum_k cl2m[p][k] * q[first + k]   (ghost ring of q read as 0 when apply_bc)
// qm is written in FP32 (state layout) and as the first GEMM's operand:
// interior or full rows/cols (interior == apply_bc), pitch op_ld, zero padded
// to op_kpad columns, as op_pieces bf16 pieces or FP32 (op_pieces == 0).
struct ModalArgs {
  const float* q;
  float* qm;
  int Nx, Ny, M;
  long plane;
  int apply_bc;
  Tables t;
  void* op;
  int op_pieces;
  long op_ld, op_piece, op_batch;
  int op_kpad;
  // folded layout (fold != 0): qm and op hold, for block b = 2p + q of table
  // plane m at batch plane b * P + m, the fold sums
  //   F[k'y][k'x] = sum_{sy,sx} (-1)^(p sy + q sx) v[mirror^sy(y)][mirror^sx(x)],
  // y = fold_phys(k'y, Ny/2), mirror(y) = Ny-1-y; qm planes are hy x hx
  // (pitch hx), the operand rows k'y < hy zero padded to op_kpad columns.
  int fold, P;
  int nlmax;  // largest member layer count (0: MAX_LAYERS)
};

// Folded index k' in [0, h) -> physical index in the first half: the pairs
// (k'+1, N-2-k') first, the ghost pair (0, N-1) last.
__host__ __device__ inline int fold_phys(int k, int h) { return k == h - 1 ? 0 : k + 1; }
void launch_to_modal(const ModalArgs& a, hipStream_t st);
void launch_to_modal_group(const ModalArgs* a, int n, hipStream_t st);  // n <= MAX_GROUP batches

// Per-mode Helmholtz residual on the full ghosted array,
//   r_z = qm_z - lap_h(psi_z) + lambda_z psi_z,
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
  int Nx, Ny, P;
  float idx2, idy2;
  const float* lam;
  // folded (fold != 0): q = fold sums, psi = parity components (4P planes of
  // Ny/2 x Nx/2, pitch Nx/2); writes the fold sums q - 4 (lap_h - lam) psi of
  // the residual.  The components satisfy the Laplacian's zero exterior and
  // the mirror rule psi[h] = +-psi[h-1] (even / odd parity) at the fold line.
  int fold;
};
void launch_residual(const ResidualArgs& a, hipStream_t st);
void launch_residual_group(const ResidualArgs* a, int n, hipStream_t st);

// Mode -> layer projection of the streamfunction:
//   psi[first + l] = sum_p cm2l[l][p] * (pm[first + p] + d[first + p])   (d nullable),
// ghost ring 0 where the member's mask is set.
struct LayerArgs {
  const float* pm;
  const float* d;
  float* psi;
  int Nx, Ny, M;
  long plane;
  Tables t;
  // mass correction (homsol nullable): pm += alpha[p] * homsol[p] before the
  // projection, with alpha[p] = -sum[p] * inv_count / hmean[p] (0 where
  // hmean[p] == 0) and sum[p] the fixed-order sum of plane p's nblk FP64
  // partials (pairs, first element, from launch_plane_sums); each block forms
  // the alphas of its member, and the member's first block writes the wall
  // value psi_wall[first + l] = sum_p cm2l[l][p] * alpha[first + p]
  const float* homsol;
  const double* partial;
  int nblk;
  const double* hmean;
  double inv_count;
  float* psi_wall;
  // folded (fold != 0): pm and d hold parity components (block b = 2p + q of
  // table plane m at plane b * P + m, Ny/2 x Nx/2, pitch Nx/2); the physical
  // value at (mirror^sy(y), mirror^sx(x)) is sum_b (-1)^(p sy + q sx) pm_b.
  int fold, P;
  int nlmax;  // largest member layer count (0: MAX_LAYERS)
};
void launch_to_layer(const LayerArgs& a, hipStream_t st);
void launch_to_layer_group(const LayerArgs* a, int n, hipStream_t st);

// One QG RHS evaluation fused with the RK stage combination:
//   out = qn + c1*(qs - qn) + c2*rhs(BC(qs), psi)   (qn given)
//   out = c1*qs + c2*rhs                           (qn null)
struct StageArgs {
  const float* qs;   // stage state (raw, ghost values kept in the combination)
  const float* psi;  // streamfunction of BC(qs)
  const float* qn;   // step start state (nullable)
  float* out;
  int Nx, Ny, P;
  long plane;               // Ny*Nx
  const float* dbeta_rows;  // [M*Ny] beta_y[j+1] - beta_y[j] per member (last entry unused)
  const float* wind_rows;   // [M*Ny] wind pattern per member
  Tables t;
  const MemberParams* mp;   // per-member parameters (null: the uniform set u)
  MemberParams u;
  int biharm;               // some member has nu4 != 0
  float idx2, idy2, iarakawa;  // 1/dx^2, 1/dy^2, 1/(12 dx dy)
  float idx, idy;              // 1/dx, 1/dy (WENO)
  float c1, c2;
  int advection;
  int apply_bc;  // 1: ghost ring of qs read as 0 in the RHS (SomaxModel.build_terms)
  int tile_rows;  // Arakawa stage tile height: 8 or 16 (0: 16)
};
void launch_qg_stage(const StageArgs& a, hipStream_t st);
// n batches sharing advection, biharm and (Arakawa) tile_rows
void launch_qg_stage_group(const StageArgs* a, int n, hipStream_t st);

// Ghost ring of nl planes of Ny x Nx set to 0 (finitevolx.zero_boundaries).
void launch_zero_ghost(float* f, int nl, int Ny, int Nx, hipStream_t st);

// Geostrophic projection of a shallow-water state (ReparameterizedQG.project):
//   q_k   = zeta(u_k, v_k) - f0 / H_k * eta_k              (pv kernel)
//   psi   = Helmholtz inversion of q (the QG solve)
//   u_k   = -(psi_k[j] - psi_k[j-1]) / dy, v_k = (psi_k[i] - psi_k[i-1]) / dx
//   eta_k = f0 H_k sum_l M[k][l] psi_l,  M = Cm2l diag(eigenvalues) Cl2m
// with zeta at the X point [j, i] (north-east corner of cell [j, i]) by forward
// differences, 0 on the last row and column; psi read as 0 outside the array.
// h is the total thickness, or the deviation eta when hdev (then ho is too).
struct ProjArgs {
  const float* h;
  const float* u;
  const float* v;
  float* q;
  const float* psi;
  float* ho;
  float* uo;
  float* vo;
  int Nx, Ny, nl;
  long plane;
  float idx, idy, f0;
  float H[MAX_LAYERS];
  float M[MAX_LAYERS * MAX_LAYERS];
  int hdev;
};
void launch_project_pv(const ProjArgs& a, hipStream_t st);
void launch_project_geo(const ProjArgs& a, hipStream_t st);

// Runge-Kutta stage state out = y + sum_{j < nk} w[j] * k[j], elementwise over
// n floats, accumulated in FP64 (out may alias y).
constexpr int LINCOMB_MAX = 7;
struct LinComb {
  const float* y;
  const float* k[LINCOMB_MAX];
  double w[LINCOMB_MAX];
  int nk;
  float* out;
  long n;
};
void launch_lincomb(const LinComb& a, hipStream_t st);

// Tsitouras' 5(4) pair (Ch. Tsitouras, "Runge-Kutta pairs of order 5(4)
// satisfying only the first column simplifying assumption", Computers &
// Mathematics with Applications 62 (2011) 770-775), the default solver of
// SomaxModel.integrate (diffrax.Tsit5).  Row i of TSIT5_A holds a[i][0..i-1]
// of stage i (stage 0 is the step start); row 6 is the solution weights b
// (first same as last: stage 6 evaluates the new state).
constexpr double TSIT5_A[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774}};

}  // namespace somax

namespace somax {

// Diagnostics of the QG models on the planes of a batch (psi already solved):
//   u = -d(psi)/dy on V faces (rows 0..Ny-2), v = d(psi)/dx on U faces
//   (columns 0..Nx-2), zeta = 5-point Laplacian (interior), and per plane
//   KE = 0.5 sum_interior (uT^2 + vT^2) dx dy with uT, vT the averages of the
//   two faces of a cell, enstrophy = 0.5 sum_interior q^2 dx dy.
// Sums are FP64, reduced in a fixed order (deterministic).
struct QGDiagArgs {
  const float* q;
  const float* psi;
  float* u;     // nullable
  float* v;     // nullable
  float* zeta;  // nullable
  int Nx, Ny, P;
  long plane;
  float idx, idy, idx2, idy2;
  double area;
  double* partial;  // [P][Ny * bx][2] workspace
  double* out;      // [P][2]: KE, enstrophy (device)
};
size_t qg_diag_partial_doubles(int Nx, int Ny, int P);
void launch_qg_diag(const QGDiagArgs& a, hipStream_t st);

// Fixed-order sum of `n` per-block pairs of every plane into out[p][2] * scale.
void launch_pair_reduce(const double* partial, int n, int P, double scale, double* out, hipStream_t st);

// Mass correction of the modal streamfunction (homogeneous solution method):
// per plane FP64 block partials of the sum of (a + b) over the full array (b
// nullable); launch_to_layer turns them into the correction (LayerArgs).
struct PlaneSumArgs {
  const float* a;
  const float* b;
  long n, plane, chunk;
  double* partial;
  int P, nblk;  // planes; blocks (partials) per plane, at most 64
};
PlaneSumArgs plane_sums_args(const float* a, const float* b, int Nx, int Ny, int P, long plane, double* partial);
int launch_plane_sums(const float* a, const float* b, int Nx, int Ny, int P, long plane, double* partial,
                      hipStream_t st);  // returns the partials per plane (at most 64)
void launch_plane_sums_group(const PlaneSumArgs* a, int n, hipStream_t st);

}  // namespace somax
