// FP64 form of the multilayer / barotropic QG model: every term of the step
// in double precision, as the reference computes it on the CPU with JAX's
// x64 mode (somax/_src/models/qg/barotropic.py:113-157 and
// somax/_src/models/qg/baroclinic.py:135-195 under jax_enable_x64).
//
// The same plan description as the FP32 / BF16 path, with precision
// SOMAXHIP_PREC_FP64, selects it.  Its kernels are written for exactness
// first (this form serves the small FP64 double gyre):
//   * layer <-> mode projections, one thread per cell and all layers;
//   * the DST-I Helmholtz solve as four batched FP64 GEMMs per inversion
//     (F S_x, S_y (.), S_y F S_x / (Lam_y + Lam_x - lambda), and back): on
//     FP64 MFMA one wave per 16 x 16 tile up to 256 x 256, LDS-tiled FMAs
//     beyond; the sine matrices built on the host with the angle reduced
//     exactly in integers first;
//   * the mass constraint's means as fixed-order block reductions;
//   * one stage kernel per RK stage: Arakawa Jacobian or WENO-Z flux form,
//     wind, bottom drag, harmonic and biharmonic viscosity and the stage
//     combination, each written in the order of operations of the
//     reference's array expressions, with floating-point contraction off so
//     that no product is fused into a sum unless written as fma().


#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "somaxhip.h"
#include "common.h"
#include "kernels.h"

#pragma clang fp contract(off)

namespace somax {

struct Qg64 {
  int nx = 0, ny = 0, Nx = 0, Ny = 0, nl = 0;
  long plane = 0;
  double dx = 0.0, dy = 0.0;
  int advection = 0, zero_psi_ghost = 0, mass = 0;
  int direct = 0;  // one layer, identity projections, no mass constraint: the transforms read q and write psi
  std::vector<void*> mem;
  // device tables
  double *sx = nullptr, *sy = nullptr;      // orthonormal DST-I matrices, Nx x Nx and Ny x Ny
  double *lamx = nullptr, *lamy = nullptr;  // their eigenvalues (2 cos(pi k/(N+1)) - 2) / d^2
  double *cl2m = nullptr, *cm2l = nullptr, *lambdas = nullptr;
  double *beta_row = nullptr, *wind_row = nullptr, *wind_scale = nullptr;
  double *homsol = nullptr, *hmean = nullptr;  // mass constraint: nl planes and their means
  // workspace layout (byte offsets)
  size_t off_A = 0, off_B = 0, off_PSI = 0, off_Q1 = 0, off_Q2 = 0, off_ALPHA = 0, ws_bytes = 0;
  ~Qg64() {
    for (void* p : mem) (void)hipFree(p);
  }
  template <class T>
  hipError_t upload(T** out, const std::vector<T>& v) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, v.size() * sizeof(T) > 0 ? v.size() * sizeof(T) : 16);
    if (e != hipSuccess) return e;
    mem.push_back(p);
    *out = static_cast<T*>(p);
    if (v.empty()) return hipSuccess;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
};

namespace {

constexpr int TILE = 64;  // GEMM output tile edge
constexpr int KT = 16;    // GEMM k-tile

// C[b] = A[b] (M x K) * B[b] (K x N), row-major, batch strides in elements
// (0: one matrix shared by the batch).  With ly set, the epilogue divides
// by (ly[row] + lx[col]) - lam[b]: the Helmholtz symbol of the solve.
struct Gemm64 {
  const double* A;
  long sA;
  int lda;
  const double* B;
  long sB;
  int ldb;
  double* C;
  long sC;
  int ldc;
  int M, N, K;
  const double* ly;
  const double* lx;
  const double* lam;
  int a_ring;  // read the outer ring of A (rows 0, M-1; columns 0, K-1) as 0
  int c_ring;  // write 0 on the outer ring of C (rows 0, M-1; columns 0, N-1)
};

// 64 x 64 output tile per workgroup of 256 threads, each thread 4 x 4
// outputs strided by 16 so the LDS reads of a wave are consecutive; the sum
// over k runs in order with one fma per product.
__global__ __launch_bounds__(256) void gemm64_kernel(Gemm64 g) {
  __shared__ double As[KT][TILE + 1];  // A tile, k-major
  __shared__ double Bs[KT][TILE + 1];
  const int b = blockIdx.z;
  const int row0 = blockIdx.y * TILE, col0 = blockIdx.x * TILE;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const double* A = g.A + b * g.sA;
  const double* B = g.B + b * g.sB;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < g.K; k0 += KT) {
    for (int e = threadIdx.x; e < TILE * KT; e += 256) {
      const int r = e / KT, k = e % KT;
      const int gr = row0 + r, gk = k0 + k;
      const bool ring = g.a_ring && (gr == 0 || gr == g.M - 1 || gk == 0 || gk == g.K - 1);
      As[k][r] = (gr < g.M && gk < g.K && !ring) ? A[(long)gr * g.lda + gk] : 0.0;
    }
    for (int e = threadIdx.x; e < KT * TILE; e += 256) {
      const int k = e / TILE, c = e % TILE;
      const int gk = k0 + k, gc = col0 + c;
      Bs[k][c] = (gk < g.K && gc < g.N) ? B[(long)gk * g.ldb + gc] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < KT; ++k) {
      double a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = Bs[k][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], bb[j], acc[i][j]);
    }
    __syncthreads();
  }
  double* C = g.C + b * g.sC;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = row0 + ty + 16 * i;
    if (row >= g.M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = col0 + tx + 16 * j;
      if (col >= g.N) continue;
      double v = acc[i][j];
      if (g.ly) v = v / ((g.ly[row] + g.lx[col]) - g.lam[b]);
      if (g.c_ring && (row == 0 || row == g.M - 1 || col == 0 || col == g.N - 1)) v = 0.0;
      C[(long)row * g.ldc + col] = v;
    }
  }
}

typedef double double4m __attribute__((ext_vector_type(4)));

// Small transforms (M, N <= 256): one wave per 16 x 16 output tile on FP64
// MFMA (v_mfma_f64_16x16x4f64, lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16] of each k-step of 4; the result row l / 16 + 4 r, column
// l % 16), the operands read from global memory (they stay in L2 at these
// sizes) GK k-steps ahead of the MFMAs.
constexpr int GK = 8;
constexpr int GEMM64_MFMA_MAX = 256;
__global__ __launch_bounds__(256) void gemm64_mfma_kernel(Gemm64 g) {
  const int tn = (g.N + 15) / 16, tiles = ((g.M + 15) / 16) * tn;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= tiles) return;
  const int lane = threadIdx.x & 63, b = blockIdx.y;
  const int i0 = (wave / tn) * 16, j0 = (wave % tn) * 16;
  const double* A = g.A + b * g.sA;
  const double* B = g.B + b * g.sB;
  const int ai = i0 + (lane & 15), bj = j0 + (lane & 15), kq = lane >> 4;
  const bool arow = ai < g.M, bcol = bj < g.N;
  const int nks = (g.K + 3) / 4;  // k-steps
  double xa[GK], xb[GK], ya[GK], yb[GK];
  auto load = [&](int s0, double (&pa)[GK], double (&pb)[GK]) {
#pragma unroll
    for (int t = 0; t < GK; ++t) {
      const int k = (s0 + t) * 4 + kq;
      const bool ring = g.a_ring && (ai == 0 || ai == g.M - 1 || k == 0 || k == g.K - 1);
      pa[t] = (arow && k < g.K && !ring) ? A[(long)ai * g.lda + k] : 0.0;
      pb[t] = (bcol && k < g.K) ? B[(long)k * g.ldb + bj] : 0.0;
    }
  };
  double4m acc = {0.0, 0.0, 0.0, 0.0};
  load(0, xa, xb);
  for (int s0 = 0; s0 < nks; s0 += GK) {
    if (s0 + GK < nks) load(s0 + GK, ya, yb);
#pragma unroll
    for (int t = 0; t < GK; ++t) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[t], xb[t], acc, 0, 0, 0);
#pragma unroll
    for (int t = 0; t < GK; ++t) {
      xa[t] = ya[t];
      xb[t] = yb[t];
    }
  }
  if (!bcol) return;
  double* C = g.C + b * g.sC;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = i0 + kq + 4 * r;
    if (row >= g.M) continue;
    double v = acc[r];
    if (g.ly) v = v / ((g.ly[row] + g.lx[bj]) - g.lam[b]);
    if (g.c_ring && (row == 0 || row == g.M - 1 || bj == 0 || bj == g.N - 1)) v = 0.0;
    C[(long)row * g.ldc + bj] = v;
  }
}

// out[l] = sum_m C[l][m] in'[m] per cell, in'[m] = in[m] (0 on the ghost ring
// when zero_in_ghost) + alpha[m] homsol[m] (when homsol is set); the ghost
// ring of out is 0 when zero_out_ghost.  One thread per cell, all layers.
struct Mix64 {
  const double* in;
  double* out;
  const double* C;
  const double* homsol;
  const double* alpha;
  int nl, Nx, Ny;
  long plane;
  int zero_in_ghost, zero_out_ghost;
};

__global__ __launch_bounds__(256) void mix64_kernel(Mix64 a) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  if (c >= a.plane) return;
  const int j = (int)(c / a.Nx), i = (int)(c % a.Nx);
  const bool ghost = j == 0 || j == a.Ny - 1 || i == 0 || i == a.Nx - 1;
  double v[MAX_LAYERS];
  for (int m = 0; m < a.nl; ++m) {
    double x = (a.zero_in_ghost && ghost) ? 0.0 : a.in[m * a.plane + c];
    if (a.homsol) x = x + a.alpha[m] * a.homsol[m * a.plane + c];
    v[m] = x;
  }
  for (int l = 0; l < a.nl; ++l) {
    double s = 0.0;
    if (!(a.zero_out_ghost && ghost)) {
      s = a.C[l * a.nl] * v[0];
      for (int m = 1; m < a.nl; ++m) s = s + a.C[l * a.nl + m] * v[m];
    }
    a.out[l * a.plane + c] = s;
  }
}

// alpha[p] = -mean(x[p]) / hmean[p], one workgroup per plane: each thread
// sums a fixed stride of the plane, then a fixed-order tree.
__global__ __launch_bounds__(256) void alpha64_kernel(const double* x, long plane, const double* hmean,
                                                      double* alpha) {
  __shared__ double red[256];
  const double* xp = x + blockIdx.x * plane;
  double s = 0.0;
  for (long c = threadIdx.x; c < plane; c += 256) s = s + xp[c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) alpha[blockIdx.x] = -(red[0] / (double)plane) / hmean[blockIdx.x];
}

__global__ __launch_bounds__(256) void fill64_kernel(double* x, long plane, const double* value) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  if (c < plane) x[blockIdx.y * plane + c] = value[blockIdx.y];
}

__global__ __launch_bounds__(256) void add_one64_kernel(double* x, long n) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  if (c < n) x[c] = 1.0 + x[c];
}

// ---------------------------------------------------------------------------
// stage kernel
// ---------------------------------------------------------------------------
struct Stage64 {
  const double* in;   // stage state (its ghost ring is read as 0 when apply_bc)
  const double* psi;  // streamfunction of the BC-applied stage state
  const double* qn;   // nullable: out = a qn / adiv + b (in + h rhs)
  double* out;
  int nl, Nx, Ny;
  long plane;
  const double* beta_row;
  const double* wind_row;
  const double* wind_scale;
  const double* cm2l;   // with alpha: wall psi of layer l = sum_m cm2l[l][m] alpha[m]
  const double* alpha;  // nullable (no mass constraint: wall psi 0)
  double nu, nu4, kappa, tau0;
  double dx, dy;
  int advection, apply_bc, rhs_only;
  double a, adiv, b, h;
};

constexpr double WENO_EPS = 1e-20;

__device__ inline double sq(double x) { return x * x; }

// WENO-Z (Borges et al. 2008) face value, upwind side first.
__device__ inline double weno5z(double qm2, double qm1, double q0, double qp1, double qp2) {
  const double p0 = (2.0 * qm2 - 7.0 * qm1 + 11.0 * q0) / 6.0;
  const double p1 = (-qm1 + 5.0 * q0 + 2.0 * qp1) / 6.0;
  const double p2 = (2.0 * q0 + 5.0 * qp1 - qp2) / 6.0;
  const double b0 = 13.0 / 12.0 * sq(qm2 - 2.0 * qm1 + q0) + 0.25 * sq(qm2 - 4.0 * qm1 + 3.0 * q0);
  const double b1 = 13.0 / 12.0 * sq(qm1 - 2.0 * q0 + qp1) + 0.25 * sq(qm1 - qp1);
  const double b2 = 13.0 / 12.0 * sq(q0 - 2.0 * qp1 + qp2) + 0.25 * sq(3.0 * q0 - 4.0 * qp1 + qp2);
  const double tau = fabs(b0 - b2);
  const double a0 = 0.1 * (1.0 + tau / (b0 + WENO_EPS));
  const double a1 = 0.6 * (1.0 + tau / (b1 + WENO_EPS));
  const double a2 = 0.3 * (1.0 + tau / (b2 + WENO_EPS));
  return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
}

__device__ inline double weno3z(double qm1, double q0, double qp1) {
  const double p0 = (-qm1 + 3.0 * q0) / 2.0;
  const double p1 = (q0 + qp1) / 2.0;
  const double b0 = sq(q0 - qm1);
  const double b1 = sq(qp1 - q0);
  const double tau = fabs(b0 - b1);
  const double a0 = (1.0 / 3.0) * (1.0 + tau / (b0 + WENO_EPS));
  const double a1 = (2.0 / 3.0) * (1.0 + tau / (b1 + WENO_EPS));
  return (a0 * p0 + a1 * p1) / (a0 + a1);
}

// Face value between interior cells f and f + 1 of a line of n interior
// cells (1..n), read through Q(c): WENO5 where its five upwind-biased cells
// are interior, else WENO3 where three are, else first-order upwind.
template <class QF>
__device__ inline double face_value(QF Q, int f, int n, double vel) {
  if (vel > 0.0) {
    if (f - 2 >= 1 && f + 2 <= n) return weno5z(Q(f - 2), Q(f - 1), Q(f), Q(f + 1), Q(f + 2));
    if (f - 1 >= 1 && f + 1 <= n) return weno3z(Q(f - 1), Q(f), Q(f + 1));
    return Q(f);
  }
  const int g = f + 1;
  if (g - 2 >= 1 && g + 2 <= n) return weno5z(Q(f + 3), Q(f + 2), Q(f + 1), Q(f), Q(f - 1));
  if (g - 1 >= 1 && g + 1 <= n) return weno3z(Q(f + 2), Q(f + 1), Q(f));
  return Q(f + 1);
}

// The 5-point Laplacian of F at (jj, ii), in the reference's order of operations.
template <class F>
__device__ inline double lap64(const Stage64& s, F f, int jj, int ii) {
  const double dx2 = s.dx * s.dx, dy2 = s.dy * s.dy;
  const double cc = f(jj, ii);
  return (f(jj, ii + 1) - 2.0 * cc + f(jj, ii - 1)) / dx2 + (f(jj + 1, ii) - 2.0 * cc + f(jj - 1, ii)) / dy2;
}

// RHS at point (j, i) of layer l, from accessors of the stage state with
// the boundary condition (Q), of the streamfunction (P), of the total PV
// q + beta (y - y0) (QT) and of the Laplacian of Q with a zero ghost ring
// (L, read by the biharmonic term); pw is the layer's wall value of psi
// (WENO-Z).
template <class QF, class PF, class QTF, class LF>
__device__ inline double rhs_point(const Stage64& s, QF Q, PF P, QTF QT, LF L, int l, int j, int i, double pw) {
  const int Nx = s.Nx, Ny = s.Ny;
  auto ghost = [&](int jj, int ii) { return jj == 0 || jj == Ny - 1 || ii == 0 || ii == Nx - 1; };
  auto lap = [&](auto F, int jj, int ii) { return lap64(s, F, jj, ii); };
  double r = 0.0;
  if (!ghost(j, i)) {
    if (s.advection == SOMAXHIP_ADV_WENO) {
      const int n_x = Nx - 2, n_y = Ny - 2;
      // corner streamfunction: corner (b, a) at T-index (a + 1/2, b + 1/2)
      auto CS = [&](int b, int a) {
        if (b == 0 || b == n_y || a == 0 || a == n_x) return pw;
        return 0.25 * (P(b, a) + P(b, a + 1) + P(b + 1, a) + P(b + 1, a + 1));
      };
      // flux through the x-face east of interior cell f of row jj
      auto xflux = [&](int jj, int f) {
        if (f == 0 || f == n_x) return 0.0;
        const double u = -(CS(jj, f) - CS(jj - 1, f)) / s.dy;
        return u * face_value([&](int cc) { return QT(jj, cc); }, f, n_x, u);
      };
      // flux through the y-face north of interior cell g of column ii
      auto yflux = [&](int g, int ii) {
        if (g == 0 || g == n_y) return 0.0;
        const double v = (CS(g, ii) - CS(g, ii - 1)) / s.dx;
        return v * face_value([&](int cc) { return QT(cc, ii); }, g, n_y, v);
      };
      r = -(xflux(j, i) - xflux(j, i - 1)) / s.dx - (yflux(j, i) - yflux(j - 1, i)) / s.dy;
    } else {
      const double pE = P(j, i + 1), pW = P(j, i - 1), pN = P(j + 1, i), pS = P(j - 1, i);
      const double pNE = P(j + 1, i + 1), pNW = P(j + 1, i - 1), pSE = P(j - 1, i + 1), pSW = P(j - 1, i - 1);
      const double qE = QT(j, i + 1), qW = QT(j, i - 1), qN = QT(j + 1, i), qS = QT(j - 1, i);
      const double qNE = QT(j + 1, i + 1), qNW = QT(j + 1, i - 1), qSE = QT(j - 1, i + 1), qSW = QT(j - 1, i - 1);
      const double jpp = (pE - pW) * (qN - qS) - (pN - pS) * (qE - qW);
      const double jpx = pE * (qNE - qSE) - pW * (qNW - qSW) - pN * (qNE - qNW) + pS * (qSE - qSW);
      const double jxp = qN * (pNE - pNW) - qS * (pSE - pSW) - qE * (pNE - pSE) + qW * (pNW - pSW);
      r = -((jpp + jpx + jxp) / (12.0 * s.dx * s.dy));
    }
  }
  r = r + (s.tau0 * s.wind_scale[l]) * s.wind_row[j];
  if (!ghost(j, i)) {
    if (s.kappa != 0.0 && l == s.nl - 1) r = r + -s.kappa * lap(P, j, i);
    if (s.nu != 0.0) r = r + s.nu * lap(Q, j, i);
    if (s.nu4 != 0.0) r = r - s.nu4 * lap(L, j, i);  // laplacian of the laplacian
  }
  return r;
}

// The stage combination: the RHS alone, or a qn / adiv + b (x + h r).
__device__ inline double combine(const Stage64& s, double x, double qn, double r) {
  if (s.rhs_only) return r;
  const double o = s.b * (x + s.h * r);
  return s.qn ? s.a * qn / s.adiv + o : o;
}

__global__ __launch_bounds__(256) void stage64_kernel(Stage64 s) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  const int l = blockIdx.y;
  if (c >= s.plane) return;
  const int Nx = s.Nx, Ny = s.Ny;
  const int j = (int)(c / Nx), i = (int)(c % Nx);
  const double* in = s.in + l * s.plane;
  const double* psi = s.psi + l * s.plane;
  // q with the boundary condition, and q + beta (y - y0) (total PV)
  auto Q = [&](int jj, int ii) {
    const bool g = jj == 0 || jj == Ny - 1 || ii == 0 || ii == Nx - 1;
    return (s.apply_bc && g) ? 0.0 : in[(long)jj * Nx + ii];
  };
  auto QT = [&](int jj, int ii) { return Q(jj, ii) + s.beta_row[jj]; };
  auto P = [&](int jj, int ii) { return psi[(long)jj * Nx + ii]; };
  double pw = 0.0;
  if (s.alpha && s.advection == SOMAXHIP_ADV_WENO) {
    pw = s.cm2l[l * s.nl] * s.alpha[0];
    for (int m = 1; m < s.nl; ++m) pw = pw + s.cm2l[l * s.nl + m] * s.alpha[m];
  }
  // the inner Laplacian of the biharmonic term, recomputed at each neighbour
  auto L = [&](int jj, int ii) {
    const bool g = jj == 0 || jj == Ny - 1 || ii == 0 || ii == Nx - 1;
    return g ? 0.0 : lap64(s, Q, jj, ii);
  };
  const double r = rhs_point(s, Q, P, QT, L, l, j, i, pw);
  s.out[l * s.plane + c] = combine(s, in[c], s.qn ? s.qn[l * s.plane + c] : 0.0, r);
}

// out = y + h * sum_j w[j] k[j] (sum in j order)
struct Lin64 {
  const double* y;
  const double* k[LINCOMB_MAX];
  double w[LINCOMB_MAX];
  int nk;
  double h;
  long n;
  double* out;
};

__global__ __launch_bounds__(256) void lincomb64_kernel(Lin64 c) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= c.n) return;
  double s = c.w[0] * c.k[0][e];
  for (int j = 1; j < c.nk; ++j) s = s + c.w[j] * c.k[j][e];
  c.out[e] = c.y[e] + c.h * s;
}

// diagnose: u = -(psi[j+1] - psi[j]) / dy on rows 0..Ny-2, v = (psi[i+1] -
// psi[i]) / dx on columns 0..Nx-2, zeta = laplacian(psi) on the interior
__global__ __launch_bounds__(256) void diag64_kernel(const double* psi, double* u, double* v, double* zeta, int Nx,
                                                     int Ny, long plane, double dx, double dy) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  if (c >= plane) return;
  const long o = blockIdx.y * plane + c;
  const double* p = psi + blockIdx.y * plane;
  const int j = (int)(c / Nx), i = (int)(c % Nx);
  if (u) u[o] = j < Ny - 1 ? -(p[c + Nx] - p[c]) / dy : 0.0;
  if (v) v[o] = i < Nx - 1 ? (p[c + 1] - p[c]) / dx : 0.0;
  if (zeta) {
    double z = 0.0;
    if (j > 0 && j < Ny - 1 && i > 0 && i < Nx - 1) {
      const double cc = p[c];
      z = (p[c + 1] - 2.0 * cc + p[c - 1]) / (dx * dx) + (p[c + Nx] - 2.0 * cc + p[c - Nx]) / (dy * dy);
    }
    zeta[o] = z;
  }
}

// per layer: kinetic energy 0.5 sum(uT^2 + vT^2) dx dy and enstrophy
// 0.5 sum(q^2) dx dy over the interior T cells, uT and vT the averages of
// the two faces of a cell; one workgroup per layer, fixed-order sums.
__global__ __launch_bounds__(256) void energy64_kernel(const double* psi, const double* q, int Nx, int Ny, long plane,
                                                       double dx, double dy, double* out) {
  __shared__ double rk[256], re[256];
  const double* p = psi + blockIdx.x * plane;
  const double* qq = q + blockIdx.x * plane;
  const long ni = (long)(Nx - 2) * (Ny - 2);
  double sk = 0.0, se = 0.0;
  for (long e = threadIdx.x; e < ni; e += 256) {
    const int j = 1 + (int)(e / (Nx - 2)), i = 1 + (int)(e % (Nx - 2));
    const long c = (long)j * Nx + i;
    const double uN = -(p[c + Nx] - p[c]) / dy, uS = -(p[c] - p[c - Nx]) / dy;
    const double vE = (p[c + 1] - p[c]) / dx, vW = (p[c] - p[c - 1]) / dx;
    const double uT = 0.5 * (uN + uS), vT = 0.5 * (vE + vW);
    sk = sk + (uT * uT + vT * vT);
    se = se + qq[c] * qq[c];
  }
  rk[threadIdx.x] = sk;
  re[threadIdx.x] = se;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rk[threadIdx.x] = rk[threadIdx.x] + rk[threadIdx.x + w];
      re[threadIdx.x] = re[threadIdx.x] + re[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double area = dx * dy;
    out[2 * blockIdx.x] = 0.5 * rk[0] * area;
    out[2 * blockIdx.x + 1] = 0.5 * re[0] * area;
  }
}

__global__ __launch_bounds__(256) void zero_ghost64_kernel(double* f, int Nx, int Ny, long plane) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  if (c >= plane) return;
  const int j = (int)(c / Nx), i = (int)(c % Nx);
  if (j == 0 || j == Ny - 1 || i == 0 || i == Nx - 1) f[blockIdx.y * plane + c] = 0.0;
}

inline unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }
inline double* wsd(void* ws, size_t off) { return reinterpret_cast<double*>(static_cast<char*>(ws) + off); }

void gemm(const Gemm64& g, int batch, hipStream_t st) {
  if (g.M <= GEMM64_MFMA_MAX && g.N <= GEMM64_MFMA_MAX) {
    const int tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
    gemm64_mfma_kernel<<<dim3((tiles + 3) / 4, batch), 256, 0, st>>>(g);
    return;
  }
  const dim3 grid((g.N + TILE - 1) / TILE, (g.M + TILE - 1) / TILE, batch);
  gemm64_kernel<<<grid, 256, 0, st>>>(g);
}

// (nabla^2 - lambda_m) phi_m = f_m over the full arrays (ghosts included):
// phi = S_y [(S_y f S_x) / (Lam_y + Lam_x - lambda)] S_x, f in `x`, phi back
// in `x`, `t` a second buffer of the same size.
void helmholtz(const Qg64& p, double* x, double* t, hipStream_t st) {
  const int Nx = p.Nx, Ny = p.Ny;
  const long pl = p.plane;
  // t = x S_x
  gemm(Gemm64{x, pl, Nx, p.sx, 0, Nx, t, pl, Nx, Ny, Nx, Nx, nullptr, nullptr, nullptr}, p.nl, st);
  // x = (S_y t) / symbol
  gemm(Gemm64{p.sy, 0, Ny, t, pl, Nx, x, pl, Nx, Ny, Nx, Ny, p.lamy, p.lamx, p.lambdas}, p.nl, st);
  // t = x S_x, x = S_y t
  gemm(Gemm64{x, pl, Nx, p.sx, 0, Nx, t, pl, Nx, Ny, Nx, Nx, nullptr, nullptr, nullptr}, p.nl, st);
  gemm(Gemm64{p.sy, 0, Ny, t, pl, Nx, x, pl, Nx, Ny, Nx, Ny, nullptr, nullptr, nullptr}, p.nl, st);
}

// psi = _invert_pv(q) (of the BC-applied q when apply_bc); alpha (the mass
// constraint's per-mode multipliers) left in the workspace.
void invert(const Qg64& p, const double* q, double* psi, void* ws, bool apply_bc, hipStream_t st) {
  double* A = wsd(ws, p.off_A);
  double* B = wsd(ws, p.off_B);
  double* alpha = wsd(ws, p.off_ALPHA);
  if (p.direct) {
    // the projections are identities: bc(q) enters the first transform (its
    // ring read as 0) and the last one writes psi (its ring 0 when masked)
    const int Nx = p.Nx, Ny = p.Ny;
    const long pl = p.plane;
    gemm(Gemm64{q, pl, Nx, p.sx, 0, Nx, B, pl, Nx, Ny, Nx, Nx, nullptr, nullptr, nullptr, apply_bc ? 1 : 0, 0}, 1,
         st);
    gemm(Gemm64{p.sy, 0, Ny, B, pl, Nx, A, pl, Nx, Ny, Nx, Ny, p.lamy, p.lamx, p.lambdas, 0, 0}, 1, st);
    gemm(Gemm64{A, pl, Nx, p.sx, 0, Nx, B, pl, Nx, Ny, Nx, Nx, nullptr, nullptr, nullptr, 0, 0}, 1, st);
    gemm(Gemm64{p.sy, 0, Ny, B, pl, Nx, psi, pl, Nx, Ny, Nx, Ny, nullptr, nullptr, nullptr, 0, p.zero_psi_ghost}, 1,
         st);
    return;
  }
  const unsigned nb = blocks(p.plane);
  mix64_kernel<<<nb, 256, 0, st>>>(
      Mix64{q, A, p.cl2m, nullptr, nullptr, p.nl, p.Nx, p.Ny, p.plane, apply_bc ? 1 : 0, 0});
  helmholtz(p, A, B, st);
  if (p.mass) alpha64_kernel<<<p.nl, 256, 0, st>>>(A, p.plane, p.hmean, alpha);
  mix64_kernel<<<nb, 256, 0, st>>>(Mix64{A, psi, p.cm2l, p.mass ? p.homsol : nullptr, alpha, p.nl, p.Nx, p.Ny,
                                         p.plane, 0, p.zero_psi_ghost});
}

Stage64 stage_args(const Qg64& p, const double* in, const double* psi, const double* qn, double* out,
                   const somaxhip_qg_params& prm, void* ws, bool apply_bc) {
  Stage64 s{};
  s.in = in;
  s.psi = psi;
  s.qn = qn;
  s.out = out;
  s.nl = p.nl;
  s.Nx = p.Nx;
  s.Ny = p.Ny;
  s.plane = p.plane;
  s.beta_row = p.beta_row;
  s.wind_row = p.wind_row;
  s.wind_scale = p.wind_scale;
  s.cm2l = p.cm2l;
  s.alpha = p.mass ? wsd(ws, p.off_ALPHA) : nullptr;
  s.nu = prm.nu;
  s.nu4 = prm.nu4;
  s.kappa = prm.kappa;
  s.tau0 = prm.tau0;
  s.dx = p.dx;
  s.dy = p.dy;
  s.advection = p.advection;
  s.apply_bc = apply_bc ? 1 : 0;
  return s;
}

void launch_stage(const Qg64& p, const Stage64& s, hipStream_t st) {
  stage64_kernel<<<dim3(blocks(p.plane), p.nl), 256, 0, st>>>(s);
}

// dq = RHS of the BC-applied q (apply_bc) or of q as given
void rhs(const Qg64& p, const double* q, double* dq, double* psi, const somaxhip_qg_params& prm, void* ws,
         bool apply_bc, hipStream_t st) {
  invert(p, q, psi, ws, apply_bc, st);
  Stage64 s = stage_args(p, q, psi, nullptr, dq, prm, ws, apply_bc);
  s.rhs_only = 1;
  launch_stage(p, s, st);
}

// nsteps SSP-RK3 steps in place: q1 = q + dt f(q), q2 = 3/4 q + 1/4 (q1 +
// dt f(q1)), q <- q/3 + 2/3 (q2 + dt f(q2)), f = the RHS of the BC-applied
// stage state.
void rk3(const Qg64& p, double* q, double* psi_out, int nsteps, double dt, const somaxhip_qg_params& prm, void* ws,
         hipStream_t st) {
  double* psi = psi_out ? psi_out : wsd(ws, p.off_PSI);
  double* q1 = wsd(ws, p.off_Q1);
  double* q2 = wsd(ws, p.off_Q2);
  auto stage = [&](const double* in, const double* qn, double* out, double a, double adiv, double b) {
    invert(p, in, psi, ws, true, st);
    Stage64 s = stage_args(p, in, psi, qn, out, prm, ws, true);
    s.a = a;
    s.adiv = adiv;
    s.b = b;
    s.h = dt;
    launch_stage(p, s, st);
  };
  for (int n = 0; n < nsteps; ++n) {
    stage(q, nullptr, q1, 0.0, 1.0, 1.0);
    stage(q1, q, q2, 0.75, 1.0, 0.25);
    stage(q2, q, q, 1.0, 3.0, 2.0 / 3.0);
  }
}

// Tsit5 with first-same-as-last: tws holds Y and K0..K6, each nl planes.
void tsit5(const Qg64& p, double* q, int nsteps, double dt, const somaxhip_qg_params& prm, void* ws, void* tws,
           hipStream_t st) {
  const long n = (long)p.nl * p.plane;
  double* psi = wsd(ws, p.off_PSI);
  double* Y = static_cast<double*>(tws);
  double* K[7];
  for (int i = 0; i < 7; ++i) K[i] = Y + (i + 1) * n;
  if (nsteps > 0) rhs(p, q, K[0], psi, prm, ws, true, st);
  for (int s = 0; s < nsteps; ++s) {
    for (int i = 1; i <= 6; ++i) {
      Lin64 c{};
      c.y = q;
      c.nk = i;
      for (int j = 0; j < i; ++j) {
        c.k[j] = K[j];
        c.w[j] = TSIT5_A[i][j];
      }
      c.h = dt;
      c.n = n;
      c.out = i < 6 ? Y : q;  // row 6 = the solution weights: the new state, in place
      lincomb64_kernel<<<blocks(n), 256, 0, st>>>(c);
      rhs(p, c.out, K[i], psi, prm, ws, true, st);
    }
    std::swap(K[0], K[6]);
  }
}

long round_up(long v, long m) { return (v + m - 1) / m * m; }

// orthonormal DST-I matrix S[n][k] = sqrt(2/(N+1)) sin(pi (n+1)(k+1)/(N+1)),
// the angle reduced modulo 2 pi in integers before the sine
std::vector<double> sine_matrix(int N) {
  std::vector<double> s((size_t)N * N);
  const double c = std::sqrt(2.0 / (N + 1));
  const long period = 2L * (N + 1);
  for (int n = 0; n < N; ++n)
    for (int k = 0; k < N; ++k) {
      const long m = ((long)(n + 1) * (k + 1)) % period;
      s[(size_t)n * N + k] = c * std::sin(M_PI * (double)m / (double)(N + 1));
    }
  return s;
}

std::vector<double> sine_eigenvalues(int N, double d) {
  std::vector<double> l(N);
  for (int k = 1; k <= N; ++k) l[k - 1] = (2.0 * std::cos(M_PI * k / (double)(N + 1)) - 2.0) / (d * d);
  return l;
}

}  // namespace

// Builds the FP64 tables of one model; null with *err set on failure.
Qg64* qg64_create(const somaxhip_qg_desc* d, std::string* err) {
  std::unique_ptr<Qg64> p(new (std::nothrow) Qg64());
  if (!p) {
    *err = "out of host memory";
    return nullptr;
  }
  p->nx = d->nx;
  p->ny = d->ny;
  p->Nx = d->nx + 2;
  p->Ny = d->ny + 2;
  p->nl = d->nl;
  p->plane = (long)p->Nx * p->Ny;
  p->dx = d->dx;
  p->dy = d->dy;
  p->advection = d->advection;
  p->zero_psi_ghost = d->zero_psi_ghost ? 1 : 0;
  p->mass = d->mass_correction ? 1 : 0;
  const int nl = d->nl;
  std::vector<double> cl2m(d->cl2m, d->cl2m + nl * nl), cm2l(d->cm2l, d->cm2l + nl * nl);
  std::vector<double> lam(d->lambdas, d->lambdas + nl), ws(d->wind_scale, d->wind_scale + nl);
  std::vector<double> beta(d->beta_y_row, d->beta_y_row + p->Ny), wind(d->wind_row, d->wind_row + p->Ny);
  hipError_t e;
  if ((e = p->upload(&p->sx, sine_matrix(p->Nx))) != hipSuccess ||
      (e = p->upload(&p->sy, sine_matrix(p->Ny))) != hipSuccess ||
      (e = p->upload(&p->lamx, sine_eigenvalues(p->Nx, p->dx))) != hipSuccess ||
      (e = p->upload(&p->lamy, sine_eigenvalues(p->Ny, p->dy))) != hipSuccess ||
      (e = p->upload(&p->cl2m, cl2m)) != hipSuccess || (e = p->upload(&p->cm2l, cm2l)) != hipSuccess ||
      (e = p->upload(&p->lambdas, lam)) != hipSuccess || (e = p->upload(&p->wind_scale, ws)) != hipSuccess ||
      (e = p->upload(&p->beta_row, beta)) != hipSuccess || (e = p->upload(&p->wind_row, wind)) != hipSuccess) {
    *err = hipGetErrorString(e);
    return nullptr;
  }
  const size_t fields = (size_t)nl * p->plane * sizeof(double);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (size_t)round_up((long)bytes, 256);
    return o;
  };
  p->off_A = take(fields);
  p->off_B = take(fields);
  p->off_PSI = take(fields);
  p->off_Q1 = take(fields);
  p->off_Q2 = take(fields);
  p->off_ALPHA = take((size_t)MAX_LAYERS * sizeof(double));
  p->ws_bytes = off;
  p->direct = nl == 1 && !p->mass && cl2m[0] == 1.0 && cm2l[0] == 1.0;
  if (p->mass) {
    // homogeneous solutions h_m = 1 + (nabla^2 - lambda_m)^-1 [lambda_m], solved
    // here by the same FP64 transforms; their means come from the description
    std::vector<double> hm(d->homsol_mean, d->homsol_mean + nl);
    double *h = nullptr, *t = nullptr;
    if ((e = p->upload(&p->hmean, hm)) != hipSuccess ||
        (e = p->upload(&p->homsol, std::vector<double>((size_t)nl * p->plane))) != hipSuccess ||
        (e = hipMalloc(&t, fields)) != hipSuccess) {
      *err = hipGetErrorString(e);
      return nullptr;
    }
    h = p->homsol;
    fill64_kernel<<<dim3(blocks(p->plane), nl), 256>>>(h, p->plane, p->lambdas);
    helmholtz(*p, h, t, nullptr);
    add_one64_kernel<<<blocks((long)nl * p->plane), 256>>>(h, (long)nl * p->plane);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(t);
    if (e != hipSuccess) {
      *err = hipGetErrorString(e);
      return nullptr;
    }
  }
  return p.release();
}

void qg64_destroy(Qg64* p) { delete p; }
size_t qg64_workspace_bytes(const Qg64* p) { return p->ws_bytes; }
size_t qg64_tsit5_workspace_bytes(const Qg64* p) { return (size_t)8 * p->nl * p->plane * sizeof(double); }

}  // namespace somax

using namespace somax;

namespace {

// The FP64 tables of an FP64 plan; sets the error and returns null otherwise.
const Qg64* f64_plan(somaxhip_qg_plan p, const void* a, const void* b, void* ws, const char* who) {
  if (!p || !a || !b || !ws) {
    set_error(std::string(who) + ": null plan, buffer or workspace");
    return nullptr;
  }
  const Qg64* q = qg_plan_f64(p);
  if (!q) set_error(std::string(who) + ": the plan was not created with SOMAXHIP_PREC_FP64");
  return q;
}

}  // namespace

extern "C" {

int somaxhip_qg_invert_f64(somaxhip_qg_plan plan, const double* q, double* psi, int apply_bc, void* ws,
                           somaxhip_stream_t stream) {
  const Qg64* p = f64_plan(plan, q, psi, ws, "somaxhip_qg_invert_f64");
  if (!p) return SOMAXHIP_ERR_ARG;
  invert(*p, q, psi, ws, apply_bc != 0, reinterpret_cast<hipStream_t>(stream));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_rhs_f64(somaxhip_qg_plan plan, const double* q, double* dq, double* psi_out,
                        const somaxhip_qg_params* prm, int apply_bc, void* ws, somaxhip_stream_t stream) {
  const Qg64* p = f64_plan(plan, q, dq, ws, "somaxhip_qg_rhs_f64");
  if (!p) return SOMAXHIP_ERR_ARG;
  if (!prm) {
    set_error("somaxhip_qg_rhs_f64: null params");
    return SOMAXHIP_ERR_ARG;
  }
  double* psi = psi_out ? psi_out : wsd(ws, p->off_PSI);
  rhs(*p, q, dq, psi, *prm, ws, apply_bc != 0, reinterpret_cast<hipStream_t>(stream));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_step_f64(somaxhip_qg_plan plan, double* q, double* psi_out, int nsteps, double dt,
                         const somaxhip_qg_params* prm, void* ws, somaxhip_stream_t stream) {
  const Qg64* p = f64_plan(plan, q, q, ws, "somaxhip_qg_step_f64");
  if (!p) return SOMAXHIP_ERR_ARG;
  if (!prm || nsteps < 0) {
    set_error("somaxhip_qg_step_f64: null params or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  rk3(*p, q, psi_out, nsteps, dt, *prm, ws, reinterpret_cast<hipStream_t>(stream));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_step_tsit5_f64(somaxhip_qg_plan plan, double* q, int nsteps, double dt,
                               const somaxhip_qg_params* prm, void* ws, void* tws, somaxhip_stream_t stream) {
  const Qg64* p = f64_plan(plan, q, q, ws, "somaxhip_qg_step_tsit5_f64");
  if (!p) return SOMAXHIP_ERR_ARG;
  if (!prm || !tws || nsteps < 0) {
    set_error("somaxhip_qg_step_tsit5_f64: null params or tsit5 workspace, or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  tsit5(*p, q, nsteps, dt, *prm, ws, tws, reinterpret_cast<hipStream_t>(stream));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_diagnose_f64(somaxhip_qg_plan plan, const double* q, double* psi, double* u, double* v, double* zeta,
                             double* ke_enstrophy, void* ws, somaxhip_stream_t stream) {
  const Qg64* p = f64_plan(plan, q, psi, ws, "somaxhip_qg_diagnose_f64");
  if (!p) return SOMAXHIP_ERR_ARG;
  if (!ke_enstrophy) {
    set_error("somaxhip_qg_diagnose_f64: null ke_enstrophy");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  invert(*p, q, psi, ws, false, st);
  if (u || v || zeta)
    diag64_kernel<<<dim3(blocks(p->plane), p->nl), 256, 0, st>>>(psi, u, v, zeta, p->Nx, p->Ny, p->plane, p->dx,
                                                                  p->dy);
  energy64_kernel<<<p->nl, 256, 0, st>>>(psi, q, p->Nx, p->Ny, p->plane, p->dx, p->dy, ke_enstrophy);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_zero_ghost_f64(double* field, int nl, int ny, int nx, somaxhip_stream_t stream) {
  if (!field || nl < 1 || nx < 1 || ny < 1) {
    set_error("somaxhip_zero_ghost_f64: null field or empty shape");
    return SOMAXHIP_ERR_ARG;
  }
  const int Nx = nx + 2, Ny = ny + 2;
  const long plane = (long)Nx * Ny;
  zero_ghost64_kernel<<<dim3(blocks(plane), nl), 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(field, Nx, Ny,
                                                                                                  plane);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

}  // extern "C"
