// Dense sine-transform GEMMs on MFMA (gfx950).
//
// Every DST-I product of the Helmholtz inversion is written as
//     O[a][b] = sum_k A[a][k] * B[b][k]
// where one operand is a data plane X (rows = data rows, k contiguous) and the
// other the symmetric sine matrix S (rows = modes, k contiguous).  With X in
// the A slot the output is X*S; with S in the A slot it is (X*S)^T.  Both
// operands are therefore read row-wise; nothing is transposed through LDS.
//
// Precision is chosen per batch plane (mode):
//   FP32 kernel   v_mfma_f32_32x32x2_f32 on FP32 planes (the plain step).
//   BF16 kernel   v_mfma_f32_32x32x16_bf16 on operands split into bf16
//                 pieces x = x0 + x1 (+ x2), FP32 accumulation:
//                   NP = 1  x0*s0
//                   NP = 3  x0*s0 + x0*s1 + x1*s0                 (~2^-17)
//                   NP = 6  ... + x0*s2 + x1*s1 + x2*s0           (~2^-25)
// Producers write the operands already split, in aligned zero-padded
// layouts, so the k-loop is a plain 16-byte copy into double-buffered LDS.
//
// Block: 256 threads = 4 waves (2x2), block tile 128x128, wave tile 64x64 =
// 2x2 MFMA tiles of 32x32.

#include "common.h"
#include "kernels.h"

namespace somax {

namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK16 = 32;          // bf16 k-tile
constexpr int LD16 = BK16 + 8;    // padded LDS row (80 bytes: conflict-free b128 reads)
constexpr int TILE16 = BM * LD16; // elements per (buffer, piece) tile
constexpr int BK32 = 16;          // fp32 k-tile
constexpr int LD32 = BK32 + 4;    // padded LDS row (80 bytes)
constexpr int TILE32 = BM * LD32;

// Epilogue: write one accumulator element; columns co >= RB are zero padding.
__device__ inline void epi_store(const GemmArgs& a, long z, int ro, int co, float v, int opieces, float lamz) {
  if (co >= a.RB) v = 0.0f;
  else if (a.lamA) v = v / (a.lamA[ro] + a.lamB[co] - lamz);
  const long idx = (long)ro * a.ldo + co;
  if (opieces == 0) {
    a.Of[z * a.o_batch + idx] = v;
    return;
  }
  uint16_t* o = a.Ob + z * a.o_batch + idx;
  const uint16_t p0 = f32_to_bf16_bits(v);
  o[0] = p0;
  if (opieces > 1) {
    const float r1 = v - bf16_bits_to_f32(p0);
    const uint16_t p1 = f32_to_bf16_bits(r1);
    o[a.o_piece] = p1;
    if (opieces > 2) o[2 * a.o_piece] = f32_to_bf16_bits(r1 - bf16_bits_to_f32(p1));
  }
}

template <int NPROD>
__device__ inline void mfma_products(const bf16x8 (&fa)[3][2], const bf16x8 (&fb)[3][2], floatx16 (&acc)[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0][i], fb[0][j], acc[i][j], 0, 0, 0);
      if (NPROD >= 3) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0][i], fb[1][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1][i], fb[0][j], acc[i][j], 0, 0, 0);
      }
      if (NPROD >= 6) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0][i], fb[2][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1][i], fb[1][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[2][i], fb[0][j], acc[i][j], 0, 0, 0);
      }
    }
}

// FP64 running sum: each k-tile's FP32 MFMA partial is added to a double
// accumulator and the MFMA accumulator restarts from 0, so the FP32 chain
// per output is one k-tile long instead of the whole contraction.
__device__ inline void flush64(floatx16 (&acc)[2][2], double (&sum)[2][2][16]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        sum[i][j][r] += (double)acc[i][j][r];
        acc[i][j][r] = 0.0f;
      }
}

// Main loop of the bf16 kernel for NPROD products (block-uniform).
// lds layout: [buf 2][operand A/B][piece NPC][row 128][LD16]
template <int NPROD, bool ACC64>
__device__ inline void bf16_mainloop(const GemmArgs& a, uint16_t* lds, long z, int a0, int b0, bool x_in_a,
                                     floatx16 (&acc)[2][2], double (&sum)[2][2][16]) {
  constexpr int NPC = NPROD == 1 ? 1 : (NPROD == 3 ? 2 : 3);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const uint16_t* X = a.Xb + z * a.x_batch;
  const uint16_t* S = a.Sb;
  const int x0 = x_in_a ? a0 : b0;
  const int s0 = x_in_a ? b0 : a0;
  // staging: per operand per piece 128 rows x 32 bf16 = 512 chunks of 16 B, 2 per thread
  const int cr = tid >> 2;        // 0..63
  const int ck = (tid & 3) * 8;   // 0,8,16,24
  const int xr0 = x0 + cr < a.xrows ? x0 + cr : a.xrows - 1;  // rows past R are never stored
  const int xr1 = x0 + cr + 64 < a.xrows ? x0 + cr + 64 : a.xrows - 1;
  const uint16_t* xg0 = X + (long)xr0 * a.ldx + ck;
  const uint16_t* xg1 = X + (long)xr1 * a.ldx + ck;
  const uint16_t* sg0 = S + (long)(s0 + cr) * a.lds + ck;
  const uint16_t* sg1 = S + (long)(s0 + cr + 64) * a.lds + ck;
  const int xo = (x_in_a ? 0 : 1) * NPC * TILE16;  // operand slot of X inside a buffer
  const int so = (x_in_a ? 1 : 0) * NPC * TILE16;
  const int st0 = cr * LD16 + ck, st1 = (cr + 64) * LD16 + ck;
  u32x4 rx0[NPC], rx1[NPC], rs0[NPC], rs1[NPC];

#define SOMAX_GLOAD(k0)                                                                   \
  _Pragma("unroll") for (int p = 0; p < NPC; ++p) {                                       \
    rx0[p] = *reinterpret_cast<const u32x4*>(xg0 + p * a.x_piece + (k0));                 \
    rx1[p] = *reinterpret_cast<const u32x4*>(xg1 + p * a.x_piece + (k0));                 \
    rs0[p] = *reinterpret_cast<const u32x4*>(sg0 + p * a.s_piece + (k0));                 \
    rs1[p] = *reinterpret_cast<const u32x4*>(sg1 + p * a.s_piece + (k0));                 \
  }
#define SOMAX_SSTORE(buf)                                                                 \
  _Pragma("unroll") for (int p = 0; p < NPC; ++p) {                                       \
    uint16_t* tb = lds + (buf) * 2 * NPC * TILE16 + p * TILE16;                           \
    *reinterpret_cast<u32x4*>(tb + xo + st0) = rx0[p];                                    \
    *reinterpret_cast<u32x4*>(tb + xo + st1) = rx1[p];                                    \
    *reinterpret_cast<u32x4*>(tb + so + st0) = rs0[p];                                    \
    *reinterpret_cast<u32x4*>(tb + so + st1) = rs1[p];                                    \
  }

  const int nk = (a.K + BK16 - 1) / BK16;
  SOMAX_GLOAD(0)
  SOMAX_SSTORE(0)
  __syncthreads();
  const int fra = (wm * 64 + (lane & 31)) * LD16 + 8 * (lane >> 5);
  const int frb = (wn * 64 + (lane & 31)) * LD16 + 8 * (lane >> 5);
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) {
      SOMAX_GLOAD((kt + 1) * BK16)
    }
    const uint16_t* tA = lds + buf * 2 * NPC * TILE16;
    const uint16_t* tB = tA + NPC * TILE16;
#pragma unroll
    for (int kk = 0; kk < BK16; kk += 16) {
      bf16x8 fa[3][2], fb[3][2];
#pragma unroll
      for (int p = 0; p < NPC; ++p)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          fa[p][t] = *reinterpret_cast<const bf16x8*>(tA + p * TILE16 + fra + t * 32 * LD16 + kk);
          fb[p][t] = *reinterpret_cast<const bf16x8*>(tB + p * TILE16 + frb + t * 32 * LD16 + kk);
        }
      mfma_products<NPROD>(fa, fb, acc);
    }
    if (ACC64) flush64(acc, sum);
    if (kt + 1 < nk) {
      SOMAX_SSTORE(buf ^ 1)
    }
    __syncthreads();
  }
#undef SOMAX_GLOAD
#undef SOMAX_SSTORE
}

}  // namespace

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void dst_gemm_bf16_kernel(GemmArgs a) {
  extern __shared__ __attribute__((aligned(16))) uint16_t lds16[];
  const long z = blockIdx.z;
  const bool lo = z < a.z_split;
  const int np = lo ? a.np_lo : a.np_hi;
  const int opieces = lo ? a.op_lo : a.op_hi;
  const int a0 = blockIdx.y * BM;
  const int b0 = blockIdx.x * BN;
  const bool x_in_a = !a.trans;

  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  double sum[2][2][16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) sum[i][j][r] = 0.0;
  if (a.acc64) {
    if (np == 1) bf16_mainloop<1, true>(a, lds16, z, a0, b0, x_in_a, acc, sum);
    else if (np == 3) bf16_mainloop<3, true>(a, lds16, z, a0, b0, x_in_a, acc, sum);
    else bf16_mainloop<6, true>(a, lds16, z, a0, b0, x_in_a, acc, sum);
  } else {
    if (np == 1) bf16_mainloop<1, false>(a, lds16, z, a0, b0, x_in_a, acc, sum);
    else if (np == 3) bf16_mainloop<3, false>(a, lds16, z, a0, b0, x_in_a, acc, sum);
    else bf16_mainloop<6, false>(a, lds16, z, a0, b0, x_in_a, acc, sum);
  }

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const float lamz = a.lamA ? a.lam[z] : 0.0f;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ro = a0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int co = b0 + wn * 64 + j * 32 + (lane & 31);
        if (ro < a.RA && co < a.RBpad) epi_store(a, z, ro, co, (float)(sum[i][j][r] + (double)acc[i][j][r]), opieces, lamz);
      }
}

// ---------------------------------------------------------------------------
// FP32 kernel (plain step)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void dst_gemm_f32_kernel(GemmArgs a) {
  __shared__ __attribute__((aligned(16))) float lds[2 * 2 * TILE32];
  const long z = blockIdx.z;
  const int a0 = blockIdx.y * BM;
  const int b0 = blockIdx.x * BN;
  const bool x_in_a = !a.trans;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const float* X = a.Xf + z * a.x_batch;
  const float* S = a.Sf;
  const int x0 = x_in_a ? a0 : b0;
  const int s0 = x_in_a ? b0 : a0;
  // staging: 128 rows x 16 floats = 512 float4, 2 per thread per operand
  const int cr = tid >> 2;
  const int ck = (tid & 3) * 4;
  const int xr0 = x0 + cr < a.xrows ? x0 + cr : a.xrows - 1;
  const int xr1 = x0 + cr + 64 < a.xrows ? x0 + cr + 64 : a.xrows - 1;
  const float* xg0 = X + (long)xr0 * a.ldx + ck;
  const float* xg1 = X + (long)xr1 * a.ldx + ck;
  const float* sg0 = S + (long)(s0 + cr) * a.lds + ck;
  const float* sg1 = S + (long)(s0 + cr + 64) * a.lds + ck;
  const int xo = x_in_a ? 0 : TILE32, so = x_in_a ? TILE32 : 0;
  const int st0 = cr * LD32 + ck, st1 = (cr + 64) * LD32 + ck;
  float4 rx0, rx1, rs0, rs1;
  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  double sum[2][2][16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) sum[i][j][r] = 0.0;
  const bool acc64 = a.acc64 != 0;
  const int nk = (a.K + BK32 - 1) / BK32;
  rx0 = *reinterpret_cast<const float4*>(xg0);
  rx1 = *reinterpret_cast<const float4*>(xg1);
  rs0 = *reinterpret_cast<const float4*>(sg0);
  rs1 = *reinterpret_cast<const float4*>(sg1);
  *reinterpret_cast<float4*>(lds + xo + st0) = rx0;
  *reinterpret_cast<float4*>(lds + xo + st1) = rx1;
  *reinterpret_cast<float4*>(lds + so + st0) = rs0;
  *reinterpret_cast<float4*>(lds + so + st1) = rs1;
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) {
      const int k1 = (kt + 1) * BK32;
      rx0 = *reinterpret_cast<const float4*>(xg0 + k1);
      rx1 = *reinterpret_cast<const float4*>(xg1 + k1);
      rs0 = *reinterpret_cast<const float4*>(sg0 + k1);
      rs1 = *reinterpret_cast<const float4*>(sg1 + k1);
    }
    const float* tA = lds + buf * 2 * TILE32;
    const float* tB = tA + TILE32;
#pragma unroll
    for (int kk = 0; kk < BK32; kk += 2) {
      float fa[2], fb[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        fa[t] = tA[(wm * 64 + t * 32 + (lane & 31)) * LD32 + kk + (lane >> 5)];
        fb[t] = tB[(wn * 64 + t * 32 + (lane & 31)) * LD32 + kk + (lane >> 5)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    if (acc64) flush64(acc, sum);
    if (kt + 1 < nk) {
      float* tb = lds + (buf ^ 1) * 2 * TILE32;
      *reinterpret_cast<float4*>(tb + xo + st0) = rx0;
      *reinterpret_cast<float4*>(tb + xo + st1) = rx1;
      *reinterpret_cast<float4*>(tb + so + st0) = rs0;
      *reinterpret_cast<float4*>(tb + so + st1) = rs1;
    }
    __syncthreads();
  }

  const float lamz = a.lamA ? a.lam[z] : 0.0f;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ro = a0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int co = b0 + wn * 64 + j * 32 + (lane & 31);
        if (ro < a.RA && co < a.RBpad) epi_store(a, z, ro, co, (float)(sum[i][j][r] + (double)acc[i][j][r]), 0, lamz);
      }
}

size_t dst_gemm_lds_bytes(int max_np) {
  const int npc = max_np == 1 ? 1 : (max_np == 3 ? 2 : 3);
  return (size_t)2 * 2 * npc * TILE16 * sizeof(uint16_t);
}

void launch_dst_gemm(const GemmArgs& a, int batch, hipStream_t st) {
  dim3 grid((a.RB + BN - 1) / BN, (a.RA + BM - 1) / BM, batch);
  if (a.Xf) {
    hipLaunchKernelGGL(dst_gemm_f32_kernel, grid, dim3(256), 0, st, a);
    return;
  }
  static bool lds_attr = false;
  if (!lds_attr) {
    // up to 120 KiB of the 160 KiB LDS of a gfx950 CU (3 pieces, double buffered)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_lds_bytes(6));
    lds_attr = true;
  }
  const int max_np = a.np_lo > a.np_hi ? a.np_lo : a.np_hi;
  hipLaunchKernelGGL(dst_gemm_bf16_kernel, grid, dim3(256), dst_gemm_lds_bytes(max_np), st, a);
}

}  // namespace somax
