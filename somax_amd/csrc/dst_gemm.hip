// Dense sine-transform GEMMs on MFMA (gfx950).
//
// Every DST-I product of the Helmholtz inversion is written as
//     O[r][c] = sum_k X[r][k] * S[c][k]          (TRANS_OUT == false)
//     O[c][r] = sum_k X[r][k] * S[c][k]          (TRANS_OUT == true)
// where X is a data plane read row-wise (k contiguous) and S is a row-major
// sine matrix.  Because S is symmetric, X*S == X*S^T, and the transposed
// output is obtained by feeding S to the MFMA A slot and X to the B slot, so
// neither operand is ever transposed through LDS.
//
// Precisions:
//   GP_FP32   v_mfma_f32_32x32x2_f32, FP32 operands (the plain step).
//   GP_BF16X3 v_mfma_f32_32x32x16_bf16, operands split x = hi + lo in bf16,
//             hi*hi + hi*lo + lo*hi accumulated in one FP32 accumulator.
//   GP_BF16X1 hi*hi only (used for the refinement correction's inverse).
// X arrives in FP32 and is split while it is staged to LDS; S is pre-split.
//
// Block: 256 threads = 4 waves in a 2x2 layout, block tile 128x128, wave
// tile 64x64 = 2x2 MFMA tiles of 32x32.

#include "common.h"
#include "kernels.h"

namespace somax {

namespace {

constexpr int BM = 128;  // A-slot rows per block
constexpr int BN = 128;  // B-slot rows per block
constexpr int PAD16 = 8;  // bf16 padding per LDS row (16 bytes)

template <int PREC>
struct GemmTraits;

template <>
struct GemmTraits<GP_FP32> {
  static constexpr int BK = 16;
};
template <>
struct GemmTraits<GP_BF16X3> {
  static constexpr int BK = 32;
};
template <>
struct GemmTraits<GP_BF16X1> {
  static constexpr int BK = 32;
};

__device__ inline uint32_t pack_bf16x2(uint16_t lo16, uint16_t hi16) {
  return (uint32_t)lo16 | ((uint32_t)hi16 << 16);
}

}  // namespace

// ---------------------------------------------------------------------------
// FP32 variant
// ---------------------------------------------------------------------------
template <bool TRANS_OUT, int EPI>
__global__ __launch_bounds__(256) void dst_gemm_f32_kernel(GemmArgs a) {
  constexpr int BK = GemmTraits<GP_FP32>::BK;
  constexpr int LD = BK + 1;
  __shared__ float sA[BM * LD];
  __shared__ float sB[BN * LD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  const long z = blockIdx.z;
  const float* X = a.X + z * a.bsx;
  float* O = a.O + z * a.bso;

  // A slot / B slot row ranges
  const int a0 = blockIdx.y * BM;
  const int b0 = blockIdx.x * BN;
  // rows of X and S served by this block
  const int x0 = TRANS_OUT ? b0 : a0;
  const int s0 = TRANS_OUT ? a0 : b0;
  float* sX = TRANS_OUT ? sB : sA;
  float* sS = TRANS_OUT ? sA : sB;

  // staging registers: X 128x16 and S 128x16 floats, 8 each per thread
  float rx[8], rs[8];
  const int ld_row = tid >> 2;        // 0..63
  const int ld_k = (tid & 3) * 4;     // 0,4,8,12

  auto prefetch = [&](int k0) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int r = x0 + ld_row + 64 * p;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = k0 + ld_k + e;
        rx[p * 4 + e] = (r < a.R && k < a.K) ? X[(long)r * a.ldx + k] : 0.0f;
      }
      const int c = s0 + ld_row + 64 * p;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = k0 + ld_k + e;
        rs[p * 4 + e] = (c < a.C && k < a.K) ? a.Sf[(long)c * a.lds + k] : 0.0f;
      }
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        sX[(ld_row + 64 * p) * LD + ld_k + e] = rx[p * 4 + e];
        sS[(ld_row + 64 * p) * LD + ld_k + e] = rs[p * 4 + e];
      }
    }
  };

  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  const int nk = (a.K + BK - 1) / BK;
  prefetch(0);
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();
    stage();
    __syncthreads();
    if (kt + 1 < nk) prefetch((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      float fa[2], fb[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        fa[t] = sA[(wm * 64 + t * 32 + (lane & 31)) * LD + kk + (lane >> 5)];
        fb[t] = sB[(wn * 64 + t * 32 + (lane & 31)) * LD + kk + (lane >> 5)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
  }

  // epilogue: accumulator row = A-slot row, col = B-slot row
  const int RA = TRANS_OUT ? a.C : a.R;
  const int RB = TRANS_OUT ? a.R : a.C;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ro = a0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int co = b0 + wn * 64 + j * 32 + (lane & 31);
        if (ro < RA && co < RB) gemm_store<EPI>(a, O, z, ro, co, acc[i][j][r]);
      }
}

// ---------------------------------------------------------------------------
// BF16 split variants
// ---------------------------------------------------------------------------
template <int PREC, bool TRANS_OUT, int EPI>
__global__ __launch_bounds__(256) void dst_gemm_bf16_kernel(GemmArgs a) {
  constexpr int BK = GemmTraits<PREC>::BK;  // 32
  constexpr int LD = BK + PAD16;            // 40 bf16 = 80 bytes per row
  constexpr bool SPLIT = (PREC == GP_BF16X3);
  __shared__ __attribute__((aligned(16))) uint16_t sAh[BM * LD];
  __shared__ __attribute__((aligned(16))) uint16_t sBh[BN * LD];
  __shared__ __attribute__((aligned(16))) uint16_t sAl[SPLIT ? BM * LD : 8];
  __shared__ __attribute__((aligned(16))) uint16_t sBl[SPLIT ? BN * LD : 8];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  const long z = blockIdx.z;
  const float* X = a.X + z * a.bsx;
  float* O = a.O + z * a.bso;

  const int a0 = blockIdx.y * BM;
  const int b0 = blockIdx.x * BN;
  const int x0 = TRANS_OUT ? b0 : a0;
  const int s0 = TRANS_OUT ? a0 : b0;
  uint16_t* sXh = TRANS_OUT ? sBh : sAh;
  uint16_t* sXl = TRANS_OUT ? sBl : sAl;
  uint16_t* sSh = TRANS_OUT ? sAh : sBh;
  uint16_t* sSl = TRANS_OUT ? sAl : sBl;

  // X staging: 128 rows x 32 k floats; thread -> (row = tid>>3 + 32p, k = (tid&7)*4 .. +3)
  float rx[4][4];
  const int xr = tid >> 3;
  const int xk = (tid & 7) * 4;
  // S staging: 128 rows x 32 k bf16 = 4 chunks of 8 per row; thread -> (row = tid>>2 + 64p, chunk = tid&3)
  uint4 rsh[2], rsl[2];
  const int sr = tid >> 2;
  const int sk = (tid & 3) * 8;

  auto prefetch = [&](int k0) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int r = x0 + xr + 32 * p;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = k0 + xk + e;
        rx[p][e] = (r < a.R && k < a.K) ? X[(long)r * a.ldx + k] : 0.0f;
      }
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int c = s0 + sr + 64 * p;
      // S is zero padded to a multiple of BK columns and of BM rows.
      const long off = (long)c * a.lds + k0 + sk;
      rsh[p] = *reinterpret_cast<const uint4*>(a.Sh + off);
      if (SPLIT) rsl[p] = *reinterpret_cast<const uint4*>(a.Sl + off);
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      uint16_t h[4], l[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        h[e] = f32_to_bf16_bits(rx[p][e]);
        if (SPLIT) l[e] = f32_to_bf16_bits(rx[p][e] - bf16_bits_to_f32(h[e]));
      }
      uint2 vh = make_uint2(pack_bf16x2(h[0], h[1]), pack_bf16x2(h[2], h[3]));
      *reinterpret_cast<uint2*>(sXh + (xr + 32 * p) * LD + xk) = vh;
      if (SPLIT) {
        uint2 vl = make_uint2(pack_bf16x2(l[0], l[1]), pack_bf16x2(l[2], l[3]));
        *reinterpret_cast<uint2*>(sXl + (xr + 32 * p) * LD + xk) = vl;
      }
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      *reinterpret_cast<uint4*>(sSh + (sr + 64 * p) * LD + sk) = rsh[p];
      if (SPLIT) *reinterpret_cast<uint4*>(sSl + (sr + 64 * p) * LD + sk) = rsl[p];
    }
  };

  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  const int nk = (a.K + BK - 1) / BK;
  prefetch(0);
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();
    stage();
    __syncthreads();
    if (kt + 1 < nk) prefetch((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      bf16x8 ah[2], bh[2], al[2], bl[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int ra = (wm * 64 + t * 32 + (lane & 31)) * LD + kk + 8 * (lane >> 5);
        const int rb = (wn * 64 + t * 32 + (lane & 31)) * LD + kk + 8 * (lane >> 5);
        ah[t] = *reinterpret_cast<const bf16x8*>(sAh + ra);
        bh[t] = *reinterpret_cast<const bf16x8*>(sBh + rb);
        if (SPLIT) {
          al[t] = *reinterpret_cast<const bf16x8*>(sAl + ra);
          bl[t] = *reinterpret_cast<const bf16x8*>(sBl + rb);
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          if (SPLIT) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          }
        }
    }
  }

  const int RA = TRANS_OUT ? a.C : a.R;
  const int RB = TRANS_OUT ? a.R : a.C;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ro = a0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int co = b0 + wn * 64 + j * 32 + (lane & 31);
        if (ro < RA && co < RB) gemm_store<EPI>(a, O, z, ro, co, acc[i][j][r]);
      }
}

// ---------------------------------------------------------------------------
// host launcher
// ---------------------------------------------------------------------------
template <int PREC, bool T, int EPI>
static void launch_one(const GemmArgs& a, int batch, hipStream_t st) {
  const int RA = T ? a.C : a.R;
  const int RB = T ? a.R : a.C;
  dim3 grid((RB + BN - 1) / BN, (RA + BM - 1) / BM, batch);
  if (PREC == GP_FP32)
    hipLaunchKernelGGL((dst_gemm_f32_kernel<T, EPI>), grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((dst_gemm_bf16_kernel<PREC, T, EPI>), grid, dim3(256), 0, st, a);
}

template <int PREC>
static void launch_prec(const GemmArgs& a, bool trans, int epi, int batch, hipStream_t st) {
  if (trans) {
    if (epi == GE_STORE) launch_one<PREC, true, GE_STORE>(a, batch, st);
    else launch_one<PREC, true, GE_FINAL>(a, batch, st);
  } else {
    if (epi == GE_STORE) launch_one<PREC, false, GE_STORE>(a, batch, st);
    else launch_one<PREC, false, GE_FINAL>(a, batch, st);
  }
}

void launch_dst_gemm(const GemmArgs& a, int prec, bool trans, int epi, int batch, hipStream_t st) {
  switch (prec) {
    case GP_FP32: launch_prec<GP_FP32>(a, trans, epi, batch, st); break;
    case GP_BF16X3: launch_prec<GP_BF16X3>(a, trans, epi, batch, st); break;
    default: launch_prec<GP_BF16X1>(a, trans, epi, batch, st); break;
  }
}

}  // namespace somax
