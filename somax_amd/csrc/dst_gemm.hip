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
// layouts.  With acc64 the FP32 MFMA partial of every few k-tiles is added to
// an FP64 running sum, so the FP32 rounding chain per output stays short.
//
// Block tile 128x128 of 32x32 MFMA tiles: 16 waves (4x4) of 32x32 for the
// x1 / x3 instances, 8 waves (2x4) of 64x32 for x6, 4 waves (2x2) of 64x64
// for FP32.  Grids are 1-D
// and remapped so that each XCD works on a contiguous run of tiles (shared
// operand rows in its L2).

#include <atomic>
#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace somax {

namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK32 = 16;          // fp32 k-tile
constexpr int LD32 = BK32 + 4;    // padded LDS row (80 bytes)
constexpr int TILE32 = BM * LD32;

// TM x TN: 32 x 32 MFMA tiles per wave (along M, along N); the 128 x 128
// block has (4 / TM) x (4 / TN) waves: 4 of 64 x 64 (TM = TN = 2), 8 of
// 64 x 32 (TM = 2, TN = 1) or 16 of 32 x 32 (TM = TN = 1).
template <int TM, int TN>
__device__ inline void zero_acc(floatx16 (&acc)[TM][TN], double (&sum)[TM][TN][16]) {
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc[i][j][r] = 0.0f;
        sum[i][j][r] = 0.0;
      }
}

// Blocks are dealt to the 8 XCDs round-robin; this remap of block b of the
// T tile blocks hands each XCD a contiguous run of tiles (same plane,
// neighbouring A rows, all B tiles) so they share its L2.
__device__ inline int remap_tiles(int b, int T) {
  const int xcd = b & 7, slot = b >> 3;
  const int per = T >> 3, extra = T & 7;
  return xcd * per + (xcd < extra ? xcd : extra) + slot;
}
// Tile t -> (plane, A-tile, B-tile).
__device__ inline void tile_of(int t, int GA, int GB, int& z, int& ta, int& tb) {
  z = t / (GA * GB);
  const int rem = t - z * GA * GB;
  ta = rem / GB;
  tb = rem - ta * GB;
}

// Epilogue.  The 128 x 128 tile leaves through LDS in two halves of 64 rows
// (the waves holding that half write their accumulators, FP64 running sum
// added when SUM64), so global stores are wide and coalesced: FP32 outputs
// with lanes along a row, bf16 pieces as 16-byte vectors of 8 columns.
// `stage` needs 64 * EPI_PITCH floats of LDS; the caller's pipeline is done.
constexpr int EPI_PITCH = 132;

// Eigenvalue rows of the spectral epilogue for one batch plane (null: none).
struct Spectral {
  const float* lamA;
  const float* lamB;
  float lamz;
};

// The spectral epilogue's eigenvalues of one tile (lamA of its 128 rows, then
// lamB of its 128 columns, both clamped to the last valid index like the
// stores) are fetched by the first 256 threads before the k-loop, so their
// latency hides under it, and staged in LDS for the epilogue.
__device__ inline float prefetch_spectral(const GemmArgs& a, const Spectral& sp, int a0, int b0) {
  const int t = threadIdx.x;
  if (t < 128) return sp.lamA[a0 + t < a.RA ? a0 + t : a.RA - 1];
  if (t < 256) return sp.lamB[b0 + t - 128 < a.RB ? b0 + t - 128 : a.RB - 1];
  return 0.0f;
}

// WN: waves along N; BW: columns of the tile (128, or 64 for the pair tiles).
template <bool SUM64, int TM, int TN, int WN = 4 / TN, int BW = 128>
__device__ inline void store_tile(const GemmArgs& a, long z, int a0, int b0, const floatx16 (&acc)[TM][TN],
                                  const double (&sum)[TM][TN][16], int opieces, float* stage, const Spectral& sp,
                                  const float* lam = nullptr) {
  constexpr int WM = 4 / TM, NT = 64 * WM * WN;  // waves along M, threads
  constexpr int TPR = BW / 8;                    // threads per row of 8-column vectors
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const float lamz = sp.lamz;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __syncthreads();  // the LDS is free (pipeline drained / previous half stored)
    if ((wm * 32 * TM) / 64 == h) {  // this wave's rows lie in half h
      const int lr0 = (wm * 32 * TM) & 63;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int lr = lr0 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int lc = wn * 32 * TN + j * 32 + (lane & 31);
            stage[lr * EPI_PITCH + lc] = SUM64 ? (float)(sum[i][j][r] + (double)acc[i][j][r]) : acc[i][j][r];
          }
    }
    __syncthreads();
    const int r0 = a0 + h * 64;
    if (opieces == 0) {
      // FP32 (any pitch): 64 lanes on 64 consecutive columns
      for (int e = tid; e < 64 * BW; e += NT) {
        const int lr = e / BW, lc = e % BW;
        const int ro = r0 + lr, co = b0 + lc;
        if (ro >= a.RA || co >= a.RBpad) continue;
        float v = stage[lr * EPI_PITCH + lc];
        if (sp.lamA) v = v / (sp.lamA[ro] + sp.lamB[co < a.RB ? co : a.RB - 1] - lamz);
        a.Of[z * a.o_batch + (long)ro * a.ldo + co] = co < a.RB ? v : 0.0f;
      }
    } else {
      // bf16 pieces: 8 consecutive columns per thread, NT / TPR rows per pass.
      // RBpad is a multiple of 32 and b0 of 64, so a vector is all in or out.
      const int lc = (tid % TPR) * 8;
      const int co = b0 + lc;
#pragma unroll
      for (int k = 0; k < 64 / (NT / TPR); ++k) {
        const int lr = tid / TPR + (NT / TPR) * k;
        const int ro = r0 + lr;
        if (ro >= a.RA || co >= a.RBpad) continue;
        const floatx4 va = *reinterpret_cast<const floatx4*>(stage + lr * EPI_PITCH + lc);
        const floatx4 vb = *reinterpret_cast<const floatx4*>(stage + lr * EPI_PITCH + lc + 4);
        float v[8] = {va[0], va[1], va[2], va[3], vb[0], vb[1], vb[2], vb[3]};
        // branch free per element (columns past RB are zero padding)
        if (sp.lamA) {
          const float la = (lam ? lam[h * 64 + lr] : sp.lamA[ro]) - lamz;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int c = co + e < a.RB ? co + e : a.RB - 1;
            const float q = v[e] / (la + (lam ? lam[128 + lc + e] : sp.lamB[c]));
            v[e] = co + e < a.RB ? q : 0.0f;
          }
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = co + e < a.RB ? v[e] : 0.0f;
        }
        u32x4 p0, p1, p2;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          uint32_t w0 = 0, w1 = 0, w2 = 0;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const float x = v[2 * e + q];
            const uint16_t b0_ = bf16_rne(x);
            const float r1 = x - bf16_bits_to_f32(b0_);
            const uint16_t b1_ = bf16_rne(r1);
            const uint16_t b2_ = bf16_rne(r1 - bf16_bits_to_f32(b1_));
            w0 |= (uint32_t)b0_ << (16 * q);
            w1 |= (uint32_t)b1_ << (16 * q);
            w2 |= (uint32_t)b2_ << (16 * q);
          }
          p0[e] = w0;
          p1[e] = w1;
          p2[e] = w2;
        }
        uint16_t* o = a.Ob + z * a.o_batch + (long)ro * a.ldo + co;
        *reinterpret_cast<u32x4*>(o) = p0;
        if (opieces > 1) *reinterpret_cast<u32x4*>(o + a.o_piece) = p1;
        if (opieces > 2) *reinterpret_cast<u32x4*>(o + 2 * a.o_piece) = p2;
      }
    }
  }
}

// Where batch plane z reads its tables and sine operand (folded or plain).
struct PlaneSel {
  int m;        // table plane (mode)
  long s_off;   // element offset of the sine operand's parity variant
  Spectral sp;  // eigenvalue rows
};

__device__ inline PlaneSel select_plane(const GemmArgs& a, long z) {
  PlaneSel s;
  int blk = 0;
  s.m = (int)z;
  if (a.fold_P) {
    s.m = (int)(z % a.fold_P);
    blk = (int)(z / a.fold_P);
  }
  const int px = blk & 1, py = blk >> 1;
  s.s_off = a.fold_P ? (long)(a.s_par ? py : px) * a.s_var : 0;
  s.sp.lamA = a.lamA ? a.lamA + px * a.lam_var : nullptr;
  s.sp.lamB = a.lamB ? a.lamB + py * a.lamB_var : nullptr;
  s.sp.lamz = a.lamA ? a.lam[s.m] : 0.0f;
  return s;
}

// One output element of a launch: spectral division, then FP32 or bf16 pieces.
__device__ inline void store_one(const GemmArgs& a, long z, int r, int c, float v, int opieces, const Spectral& sp) {
  if (sp.lamA) v = v / (sp.lamA[r] + sp.lamB[c] - sp.lamz);
  const long o = z * a.o_batch + (long)r * a.ldo + c;
  if (opieces == 0) {
    a.Of[o] = v;
    return;
  }
  const uint16_t b0 = bf16_rne(v);
  a.Ob[o] = b0;
  if (opieces > 1) {
    const float r1 = v - bf16_bits_to_f32(b0);
    const uint16_t b1 = bf16_rne(r1);
    a.Ob[o + a.o_piece] = b1;
    if (opieces > 2) a.Ob[o + 2 * a.o_piece] = bf16_rne(r1 - bf16_bits_to_f32(b1));
  }
}

// Fringe workgroup `part` of `fpw` for batch plane z: the dot products of
// rows [am, RA) x columns [0, RB) and rows [0, am) x columns [bm, RB), then
// its share of the zero padding [RB, RBpad) of the rows that no tile covers.
// Every dot pairs one of the at most four fixed rows (A rows [am, RA), B rows
// [bm, RB)) with a variable row.  The fixed rows are staged once in LDS,
// reconstructed from their pieces; eight lanes share a dot along k, two dots
// per lane group are in flight, and the loads of KB 64-element k-chunks are
// issued before any is used.  Accumulation is FP64 for FP32 operands and
// six-product planes, FP32 otherwise (the precision of those products).
template <bool F32, int NPC, int NT, int KBX = 0>
__device__ void gemm_fringe(const GemmArgs& a, long z, int part, const GemmCut& cut, const PlaneSel& ps, int opieces,
                            void* lds) {
  constexpr int G = NT / 8;  // lane groups
  using acc_t = typename std::conditional<(F32 || NPC == 3), double, float>::type;
  // k-chunks per batch of loads (register bound; KBX: the caller's budget).
  // Each lane sums the same k values in the same order whatever the batch.
  constexpr int KB = KBX ? KBX : ((NPC == 3 || NT >= 1024) ? 4 : 9);
  constexpr int NV = F32 ? 2 : NPC;  // 16-byte loads per variable-row chunk
  const int fpw = cut.fpw;
  const int fa = a.RA - cut.am, fb = a.RB - cut.bm;
  const int nrow = fa * a.RB;
  const int dots = nrow + cut.am * fb;
  const int tid = threadIdx.x, g = tid >> 3, l8 = tid & 7;
  const bool x_in_a = !a.trans;
  const int esz = F32 ? 4 : 2;
  const char* X = F32 ? reinterpret_cast<const char*>(a.Xf + z * a.x_batch)
                      : reinterpret_cast<const char*>(a.Xb + z * a.x_batch);
  const char* S = F32 ? reinterpret_cast<const char*>(a.Sf + ps.s_off) : reinterpret_cast<const char*>(a.Sb + ps.s_off);
  auto xrow = [&](int r) { return X + (long)(r < a.xrows ? r : a.xrows - 1) * a.ldx * esz; };
  auto srow = [&](int r) { return S + (long)r * a.lds * esz; };
  // row of the A operand (is_a) or of the B operand, and its piece stride
  auto row_of = [&](bool is_a, int r, long& piece) {
    const bool from_x = is_a == x_in_a;
    piece = from_x ? a.x_piece : a.s_piece;
    return from_x ? xrow(r) : srow(r);
  };
  // 8 reconstructed operand values at k0 (k0 < K, 8-aligned)
  auto load_chunk = [&](const char* row, long piece, int k0, acc_t (&v)[8]) {
    if constexpr (F32) {
      const float* f = reinterpret_cast<const float*>(row) + k0;
      const floatx4 x0 = *reinterpret_cast<const floatx4*>(f);
      const floatx4 x1 = *reinterpret_cast<const floatx4*>(f + 4);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        v[e] = (acc_t)x0[e];
        v[4 + e] = (acc_t)x1[e];
      }
    } else {
      const uint16_t* b = reinterpret_cast<const uint16_t*>(row) + k0;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = (acc_t)0;
#pragma unroll
      for (int p = 0; p < NPC; ++p) {
        const u32x4 w = *reinterpret_cast<const u32x4*>(b + p * piece);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t u = w[e];
          v[2 * e] += (acc_t)__builtin_bit_cast(float, u << 16);
          v[2 * e + 1] += (acc_t)__builtin_bit_cast(float, u & 0xffff0000u);
        }
      }
    }
  };

  // zero padding first (stores need no wait): rows [am, RA), and every row
  // when the columns have a fringe; 32 lanes per row, columns RB + lane
  {
    const int padw = a.RBpad - a.RB;
    const int nzr = fb > 0 ? a.RA : fa;
    if (padw > 0 && nzr > 0) {
      const int zper = (nzr + fpw - 1) / fpw;
      const int r0 = part * zper, r1 = r0 + zper < nzr ? r0 + zper : nzr;
      for (int e = r0 * 32 + tid; e < r1 * 32; e += NT) {
        const int rq = e >> 5;
        const int r = fb > 0 ? rq : cut.am + rq;
        for (int cq = e & 31; cq < padw; cq += 32) {
          const long o = z * a.o_batch + (long)r * a.ldo + a.RB + cq;
          if (opieces == 0) {
            a.Of[o] = 0.0f;
          } else {
            for (int p = 0; p < opieces; ++p) a.Ob[o + p * a.o_piece] = 0;
          }
        }
      }
    }
  }
  // fixed rows -> LDS: fix[i][k], i < fa: A row am + i; else B row bm + i - fa
  const int K8 = (a.K + 7) & ~7;
  acc_t* fix = static_cast<acc_t*>(lds);
  const int nfix = fa + fb;
  for (int e = tid; e < nfix * (K8 / 8); e += NT) {
    const int i = e / (K8 / 8), k0 = (e - i * (K8 / 8)) * 8;
    long piece;
    const char* row = i < fa ? row_of(true, cut.am + i, piece) : row_of(false, cut.bm + i - fa, piece);
    acc_t v[8];
    load_chunk(row, piece, k0, v);
#pragma unroll
    for (int q = 0; q < 8; ++q) fix[i * K8 + k0 + q] = v[q];
  }
  __syncthreads();

  const int per = (dots + fpw - 1) / fpw;
  const int d0 = part * per;
  const int d1 = d0 + per < dots ? d0 + per : dots;
  for (int base = d0; base < d1; base += 2 * G) {
    const char* vrow[2];
    long vpiece[2];
    int fixi[2], rr[2], cc[2];
    bool live[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int d = base + G * h + g;
      live[h] = d < d1;
      int r = 0, c = 0, f = 0;
      bool var_is_b = true;
      if (live[h]) {
        if (d < nrow) {
          const int q = (unsigned)d / (unsigned)a.RB;
          r = cut.am + q;
          c = d - q * a.RB;
          f = r - cut.am;
        } else {
          const int e = d - nrow;
          const int q = fb == 1 ? e : e >> 1;  // fb is 1 or 2
          r = q;
          c = cut.bm + (e - q * fb);
          f = fa + c - cut.bm;
          var_is_b = false;
        }
      }
      rr[h] = r;
      cc[h] = c;
      fixi[h] = f;
      vrow[h] = var_is_b ? row_of(false, c, vpiece[h]) : row_of(true, r, vpiece[h]);
    }
    acc_t acc[2] = {(acc_t)0, (acc_t)0};
    for (int kb = 0; kb < a.K; kb += 64 * KB) {
      u32x4 raw[2][KB][NV];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < KB; ++j) {
          const int k0 = kb + 64 * j + l8 * 8;
#pragma unroll
          for (int w = 0; w < NV; ++w) {
            const char* src = vrow[h] + (long)k0 * esz + (F32 ? 16 * w : w * vpiece[h] * esz);
            raw[h][j][w] = (live[h] && k0 < a.K) ? *reinterpret_cast<const u32x4*>(src) : u32x4{0u, 0u, 0u, 0u};
          }
        }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < KB; ++j) {
          const int k0 = kb + 64 * j + l8 * 8;
          if (k0 >= a.K) continue;
          acc_t v[8];
          if constexpr (F32) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const uint32_t u0 = raw[h][j][0][e], u1 = raw[h][j][1][e];
              v[e] = (acc_t)__builtin_bit_cast(float, u0);
              v[4 + e] = (acc_t)__builtin_bit_cast(float, u1);
            }
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (acc_t)0;
#pragma unroll
            for (int w = 0; w < NV; ++w)
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                const uint32_t u = raw[h][j][w][e];
                v[2 * e] += (acc_t)__builtin_bit_cast(float, u << 16);
                v[2 * e + 1] += (acc_t)__builtin_bit_cast(float, u & 0xffff0000u);
              }
          }
          const acc_t* fx = fix + fixi[h] * K8 + k0;
#pragma unroll
          for (int e = 0; e < 8; ++e) acc[h] += v[e] * fx[e];
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      acc_t t = acc[h];
      t += __shfl_xor(t, 1);
      t += __shfl_xor(t, 2);
      t += __shfl_xor(t, 4);
      if (live[h] && l8 == 0) store_one(a, z, rr[h], cc[h], (float)t, opieces, ps.sp);
    }
  }
}

template <int NPROD, int TM, int TN>
__device__ inline void mfma_products(const bf16x8 (&fa)[3][TM], const bf16x8 (&fb)[3][TN], floatx16 (&acc)[TM][TN]) {
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
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

// ---------------------------------------------------------------------------
// BF16 kernel
//
// Tiles of 128 rows x 32 k per operand piece live in LDS unpadded (64 B per
// row) with the 16-byte chunk index XOR-swizzled by swz(row), which keeps the
// b128 fragment reads of 8 or 16 consecutive rows on distinct banks.  They are
// filled by global_load_lds (16 bytes per lane straight into LDS, no staging
// registers); each wave instruction fills 16 rows (1 KiB, linear, so the
// writes are conflict free).  NSTAGE buffers keep NSTAGE-1 k-tiles in flight.
// ---------------------------------------------------------------------------
__device__ inline int swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// k-tile stages in LDS per operand piece count (1, 2 or 3 bf16 pieces).  Five
// stages for two pieces (160 KiB) and eight for one were measured no faster.
__host__ __device__ constexpr int bf16_nstage(int npc) { return npc == 3 ? 3 : 4; }

template <int NPROD, int TM, int TN>
struct Bf16Cfg {
  static constexpr int NPC = NPROD == 1 ? 1 : (NPROD == 3 ? 2 : 3);
  static constexpr int NSTAGE = bf16_nstage(NPC);
  static constexpr int TILE = 128 * 32;         // bf16 elements per operand piece
  static constexpr int STAGE = 2 * NPC * TILE;  // A pieces then B pieces
  static constexpr int NWAVES = (4 / TM) * (4 / TN);
  static constexpr int OPW = NWAVES >= 16 ? 1 : 2;           // operands each wave loads
  static constexpr int RPT = 16 / (NWAVES * OPW);            // 16-row blocks per operand piece
  static constexpr int LPT = OPW * NPC * RPT;                // global_load_lds per thread per k-tile
  static constexpr int FLUSH = 2;               // k-tiles between FP64 flushes
};

template <int N>
__device__ inline void wait_vmcnt() {
  // s_waitcnt vmcnt(N); expcnt and lgkmcnt fields at their maxima (no wait)
  __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

// Wait until at most `later` k-tiles of LPT loads each are outstanding
// (later <= MAXL, a wave-uniform value).
template <int LPT, int MAXL>
__device__ inline void wait_later(int later) {
  if constexpr (MAXL <= 0) {
    wait_vmcnt<0>();
  } else {
    static_assert(MAXL * LPT < 64, "vmcnt holds at most 63");
    if (later >= MAXL) wait_vmcnt<MAXL * LPT>();
    else wait_later<LPT, MAXL - 1>(later);
  }
}

template <int NPROD, bool ACC64, int TM, int TN>
__device__ inline void bf16_mainloop(const GemmArgs& a, const uint16_t* S, uint16_t* lds, long z, int a0, int b0,
                                     bool x_in_a, floatx16 (&acc)[TM][TN], double (&sum)[TM][TN][16]) {
  using C = Bf16Cfg<NPROD, TM, TN>;
  constexpr int NPC = C::NPC;
  constexpr int RPT = C::RPT;
  constexpr int OPW = C::OPW;
  constexpr int WN = 4 / TN;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const uint16_t* X = a.Xb + z * a.x_batch;
  const int x0 = x_in_a ? a0 : b0;
  const int s0 = x_in_a ? b0 : a0;
  const int xslot = x_in_a ? 0 : NPC;  // piece-tile index of X inside a stage
  const int sslot = x_in_a ? NPC : 0;

  // this lane's RPT rows of each 128-row tile and the swizzled source chunk
  const uint16_t* xsrc[RPT];
  const uint16_t* ssrc[RPT];
  int ldst[RPT];
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    // first row filled by this wave instruction (with one operand per wave,
    // waves 0-7 load X and 8-15 load S)
    const int r0 = ((OPW == 2 ? wave : (wave & 7)) * RPT + j) * 16;
    const int row = r0 + (lane >> 2);
    const int chunk = (lane & 3) ^ swz(row);
    const int xr = x0 + row < a.xrows ? x0 + row : a.xrows - 1;  // rows past R are never stored
    xsrc[j] = X + (long)xr * a.ldx + chunk * 8;
    ssrc[j] = S + (long)(s0 + row) * a.lds + chunk * 8;
    ldst[j] = r0 * 32;  // wave-uniform LDS element offset of the 16-row block
  }

  const bool load_x = OPW == 2 || wave < 8, load_s = OPW == 2 || wave >= 8;
#define SOMAX_ISSUE(KT, STG)                                                                         \
  {                                                                                                  \
    uint16_t* sb_ = lds + (STG) * C::STAGE;                                                          \
    const int k0_ = (KT) * 32;                                                                       \
    _Pragma("unroll") for (int p = 0; p < NPC; ++p)                                                 \
      _Pragma("unroll") for (int j = 0; j < RPT; ++j) {                                              \
        if (load_x)                                                                                  \
          __builtin_amdgcn_global_load_lds(xsrc[j] + p * a.x_piece + k0_, sb_ + (xslot + p) * C::TILE + ldst[j], \
                                           16, 0, 0);                                                \
        if (load_s)                                                                                  \
          __builtin_amdgcn_global_load_lds(ssrc[j] + p * a.s_piece + k0_, sb_ + (sslot + p) * C::TILE + ldst[j], \
                                           16, 0, 0);                                                \
      }                                                                                              \
  }

  const int nk = (a.K + 31) / 32;
#pragma unroll
  for (int s = 0; s < C::NSTAGE - 1; ++s)
    if (s < nk) SOMAX_ISSUE(s, s)

  int ra[TM], rb[TN];
#pragma unroll
  for (int t = 0; t < TM; ++t) ra[t] = wm * 32 * TM + t * 32 + (lane & 31);
#pragma unroll
  for (int t = 0; t < TN; ++t) rb[t] = wn * 32 * TN + t * 32 + (lane & 31);
  // fragments of k16 step kk of the tile in stage buffer sb
  auto read_frags = [&](const uint16_t* sb, int kk, bf16x8(&fa)[3][TM], bf16x8(&fb)[3][TN]) {
    const int c = kk * 2 + (lane >> 5);  // logical 16-byte chunk of this lane's k-slice
#pragma unroll
    for (int p = 0; p < NPC; ++p) {
#pragma unroll
      for (int t = 0; t < TM; ++t)
        fa[p][t] = *reinterpret_cast<const bf16x8*>(sb + p * C::TILE + ra[t] * 32 + ((c ^ swz(ra[t])) << 3));
#pragma unroll
      for (int t = 0; t < TN; ++t)
        fb[p][t] = *reinterpret_cast<const bf16x8*>(sb + (NPC + p) * C::TILE + rb[t] * 32 + ((c ^ swz(rb[t])) << 3));
    }
  };
  // Make tile kt current: its loads (this wave's) have landed once at most
  // min(NSTAGE-2, nk-1-kt) later tiles are outstanding; after the barrier
  // every wave's loads of kt are visible and every wave has finished reading
  // tile kt-1, whose stage then takes tile kt+NSTAGE-1.  The caller has
  // waited for its own LDS reads (lgkmcnt(0)) first.  The empty asm
  // statements keep the compiler from moving LDS reads across the barrier.
  auto enter_tile = [&](int kt) {
    const int later = (nk - 1 - kt) < (C::NSTAGE - 2) ? (nk - 1 - kt) : (C::NSTAGE - 2);
    wait_later<C::LPT, C::NSTAGE - 2>(later);
    __builtin_amdgcn_s_waitcnt((0xF) | (0x7 << 4) | (0x0 << 8) | (0x3 << 14));  // lgkmcnt(0)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (kt + C::NSTAGE - 1 < nk) SOMAX_ISSUE(kt + C::NSTAGE - 1, (kt + C::NSTAGE - 1) % C::NSTAGE)
  };

  // Software pipeline: the fragments of the next k16 step are read while the
  // MFMAs of the current one run; the barrier that hands over the next tile
  // sits between the two steps of a tile, so its first fragments also load
  // under the current tile's second-step MFMAs.
  // Four fragment register sets rotate over the two k16 steps of two
  // consecutive tiles, so a read never targets registers that a queued MFMA
  // still has to consume (that write-after-read would hold the read back
  // until the MFMA starts and serialise LDS latency with MFMA execution).
  // (Two sets when two waves share a SIMD and three-piece fragments would not
  // fit: the other wave covers the read-after-MFMA wait.)
  constexpr int NS = (TN == 1 && NPC == 3) ? 2 : 4;
  bf16x8 fsa[NS][3][TM], fsb[NS][3][TN];
  auto& fa0 = fsa[0];
  auto& fb0 = fsb[0];
  auto& fa1 = fsa[1];
  auto& fb1 = fsb[1];
  auto& fa2 = fsa[2 % NS];
  auto& fb2 = fsb[2 % NS];
  auto& fa3 = fsa[3 % NS];
  auto& fb3 = fsb[3 % NS];
  auto stage_of = [&](int kt) { return lds + (kt % C::NSTAGE) * C::STAGE; };
  auto lgkm0 = [] { __builtin_amdgcn_s_waitcnt((0xF) | (0x7 << 4) | (0x0 << 8) | (0x3 << 14)); };
  enter_tile(0);
  read_frags(stage_of(0), 0, fa0, fb0);
  // One tile that hands over to its successor: step 0 from (xa, xb), step 1
  // into (ya, yb), next tile's step 0 into (za, zb).  sched_barrier(0) fences
  // keep the machine scheduler from sinking reads below the MFMAs they overlap.
#define SOMAX_TILE_MID(KT, W, XA, XB, YA, YB, ZA, ZB)       \
  {                                                         \
    lgkm0();                                                \
    read_frags(stage_of(KT), 1, YA, YB);                    \
    __builtin_amdgcn_sched_barrier(0);                      \
    mfma_products<NPROD, TM, TN>(XA, XB, W);                        \
    __builtin_amdgcn_sched_barrier(0);                      \
    enter_tile((KT) + 1);                                   \
    read_frags(stage_of((KT) + 1), 0, ZA, ZB);              \
    __builtin_amdgcn_sched_barrier(0);                      \
    mfma_products<NPROD, TM, TN>(YA, YB, W);                        \
    __builtin_amdgcn_sched_barrier(0);                      \
  }
  // two tiles, leaving the fragments of tile kt+2 step 0 in set 0
  auto pair = [&](int kt, floatx16(&w)[TM][TN]) {
    SOMAX_TILE_MID(kt, w, fa0, fb0, fa1, fb1, fa2, fb2)
    SOMAX_TILE_MID(kt + 1, w, fa2, fb2, fa3, fb3, fa0, fb0)
  };
  int kt = 0;
  if (ACC64) {
    // FP64 running sum over windows of two k-tiles: each window's MFMA chain
    // starts from zero and is added to the sum at its end, so no FP32
    // accumulator is carried around the loop (it stays in AGPRs).
    static_assert(C::FLUSH == 2, "one window per tile pair");
    for (; kt + 2 <= nk - 1; kt += 2) {
      floatx16 w[TM][TN];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) w[i][j] = floatx16{};
      pair(kt, w);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) sum[i][j][r] += (double)w[i][j][r];
    }
  } else {
    for (; kt + 2 <= nk - 1; kt += 2) pair(kt, acc);
  }
  // remaining one or two tiles (the last has no successor); set 0 holds step 0
  if (kt == nk - 2) {
    SOMAX_TILE_MID(kt, acc, fa0, fb0, fa1, fb1, fa2, fb2)
    lgkm0();
    read_frags(stage_of(nk - 1), 1, fa3, fb3);
    mfma_products<NPROD, TM, TN>(fa2, fb2, acc);
    mfma_products<NPROD, TM, TN>(fa3, fb3, acc);
  } else {
    lgkm0();
    read_frags(stage_of(nk - 1), 1, fa1, fb1);
    mfma_products<NPROD, TM, TN>(fa0, fb0, acc);
    mfma_products<NPROD, TM, TN>(fa1, fb1, acc);
  }
#undef SOMAX_TILE_MID
#undef SOMAX_ISSUE
}

// ---------------------------------------------------------------------------
// Pair tiles (x1 / x3 transforms): 128 x 64 output tiles, 8 waves of 32 x 32
// (4 along M, 2 along N), NSTAGE_P stages of a 128-row A tile and a 64-row B
// tile per operand piece, at most 72 KiB of LDS and 128 VGPRs per lane, so
// two workgroups share a CU and hide each other's k-loop latency.  Each wave
// loads 16 A rows and 8 B rows per piece (the B load with lanes 0-31), so
// every wave issues the same number of loads per k-tile.  Per output the same
// MFMA chain as the 128 x 128 tiles.
// ---------------------------------------------------------------------------
constexpr int BNP = 64;
constexpr int NSTAGE_P = 3;
template <int NPROD>
struct PairCfg {
  static constexpr int NPC = NPROD == 1 ? 1 : 2;
  static constexpr int TA = 128 * 32;                // bf16 elements per A piece tile
  static constexpr int TB = BNP * 32;                // per B piece tile
  static constexpr int STAGE = NPC * (TA + TB);
  static constexpr int LPT = 2 * NPC;                // loads per thread per k-tile
};

template <int NPROD>
__device__ inline void bf16_mainloop_pair(const GemmArgs& a, const uint16_t* S, uint16_t* lds, long z, int a0, int b0,
                                          bool x_in_a, floatx16 (&acc)[1][1]) {
  using C = PairCfg<NPROD>;
  constexpr int NPC = C::NPC;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const uint16_t* X = a.Xb + z * a.x_batch;
  // A rows a0 + ra_l (16 per wave), B rows b0 + rb_l (8 per wave, lanes 0-31)
  const int ra_l = wave * 16 + (lane >> 2);
  const int rb_l = wave * 8 + ((lane & 31) >> 2);
  const int ca = (lane & 3) ^ swz(ra_l), cb = (lane & 3) ^ swz(rb_l);
  auto row_src = [&](bool is_a, int r, int chunk, long& piece) {
    const bool from_x = is_a == x_in_a;
    if (from_x) {
      piece = a.x_piece;
      const int xr = r < a.xrows ? r : a.xrows - 1;  // rows past R are never stored
      return X + (long)xr * a.ldx + chunk * 8;
    }
    piece = a.s_piece;
    return S + (long)r * a.lds + chunk * 8;
  };
  long pa, pb;
  const uint16_t* asrc = row_src(true, a0 + ra_l, ca, pa);
  const uint16_t* bsrc = row_src(false, b0 + rb_l, cb, pb);
  const int adst = wave * 16 * 32, bdst = NPC * C::TA + wave * 8 * 32;
  const bool bload = lane < 32;
#define SOMAX_ISSUE_P(KT, STG)                                                                             \
  {                                                                                                        \
    uint16_t* sb_ = lds + (STG) * C::STAGE;                                                                \
    const int k0_ = (KT) * 32;                                                                             \
    _Pragma("unroll") for (int p = 0; p < NPC; ++p) {                                                      \
      __builtin_amdgcn_global_load_lds(asrc + p * pa + k0_, sb_ + p * C::TA + adst, 16, 0, 0);             \
      if (bload) __builtin_amdgcn_global_load_lds(bsrc + p * pb + k0_, sb_ + bdst + p * C::TB, 16, 0, 0);  \
    }                                                                                                      \
  }
  const int nk = (a.K + 31) / 32;
#pragma unroll
  for (int s = 0; s < NSTAGE_P - 1; ++s)
    if (s < nk) SOMAX_ISSUE_P(s, s)

  const int ra = wm * 32 + (lane & 31), rb = wn * 32 + (lane & 31);
  auto read_frags = [&](const uint16_t* sb, int kk, bf16x8(&fa)[3][1], bf16x8(&fb)[3][1]) {
    const int c = kk * 2 + (lane >> 5);
#pragma unroll
    for (int p = 0; p < NPC; ++p) {
      fa[p][0] = *reinterpret_cast<const bf16x8*>(sb + p * C::TA + ra * 32 + ((c ^ swz(ra)) << 3));
      fb[p][0] = *reinterpret_cast<const bf16x8*>(sb + NPC * C::TA + p * C::TB + rb * 32 + ((c ^ swz(rb)) << 3));
    }
  };
  auto enter_tile = [&](int kt) {
    const int later = (nk - 1 - kt) < (NSTAGE_P - 2) ? (nk - 1 - kt) : (NSTAGE_P - 2);
    wait_later<C::LPT, NSTAGE_P - 2>(later);
    __builtin_amdgcn_s_waitcnt((0xF) | (0x7 << 4) | (0x0 << 8) | (0x3 << 14));  // lgkmcnt(0)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (kt + NSTAGE_P - 1 < nk) SOMAX_ISSUE_P(kt + NSTAGE_P - 1, (kt + NSTAGE_P - 1) % NSTAGE_P)
  };
  bf16x8 fsa[4][3][1], fsb[4][3][1];
  auto stage_of = [&](int kt) { return lds + (kt % NSTAGE_P) * C::STAGE; };
  auto lgkm0 = [] { __builtin_amdgcn_s_waitcnt((0xF) | (0x7 << 4) | (0x0 << 8) | (0x3 << 14)); };
  enter_tile(0);
  read_frags(stage_of(0), 0, fsa[0], fsb[0]);
#define SOMAX_TILE_MID_P(KT, XS, YS, ZS)                   \
  {                                                        \
    lgkm0();                                               \
    read_frags(stage_of(KT), 1, fsa[YS], fsb[YS]);         \
    __builtin_amdgcn_sched_barrier(0);                     \
    mfma_products<NPROD, 1, 1>(fsa[XS], fsb[XS], acc);     \
    __builtin_amdgcn_sched_barrier(0);                     \
    enter_tile((KT) + 1);                                  \
    read_frags(stage_of((KT) + 1), 0, fsa[ZS], fsb[ZS]);   \
    __builtin_amdgcn_sched_barrier(0);                     \
    mfma_products<NPROD, 1, 1>(fsa[YS], fsb[YS], acc);     \
    __builtin_amdgcn_sched_barrier(0);                     \
  }
  int kt = 0;
  for (; kt + 2 <= nk - 1; kt += 2) {
    SOMAX_TILE_MID_P(kt, 0, 1, 2)
    SOMAX_TILE_MID_P(kt + 1, 2, 3, 0)
  }
  if (kt == nk - 2) {
    SOMAX_TILE_MID_P(kt, 0, 1, 2)
    lgkm0();
    read_frags(stage_of(nk - 1), 1, fsa[3], fsb[3]);
    mfma_products<NPROD, 1, 1>(fsa[2], fsb[2], acc);
    mfma_products<NPROD, 1, 1>(fsa[3], fsb[3], acc);
  } else {
    lgkm0();
    read_frags(stage_of(nk - 1), 1, fsa[1], fsb[1]);
    mfma_products<NPROD, 1, 1>(fsa[0], fsb[0], acc);
    mfma_products<NPROD, 1, 1>(fsa[1], fsb[1], acc);
  }
#undef SOMAX_TILE_MID_P
#undef SOMAX_ISSUE_P
}

}  // namespace

size_t dst_gemm_lds_bytes(int max_np) {
  const int npc = max_np == 1 ? 1 : (max_np == 3 ? 2 : 3);
  return (size_t)bf16_nstage(npc) * 2 * npc * 128 * 32 * sizeof(uint16_t);
}
size_t dst_gemm_pair_lds_bytes(int max_np) {
  const int npc = max_np == 1 ? 1 : 2;
  return (size_t)NSTAGE_P * npc * (128 + BNP) * 32 * sizeof(uint16_t);
}

// Workgroup shape per instance: x1 / x3 run 16 waves of 32 x 32 (four per
// SIMD), x6 8 waves of 64 x 32 (its three-piece fragments need the
// registers).  Several waves per SIMD let one wave's MFMAs overlap another's
// LDS traffic: 8 waves are 23-33% faster than 4 at the same tile, and 16 a
// further 12-20% on the many small tiles of ensemble grids.
template <int NPMAX>
struct KernelShape {
  static constexpr int TM = NPMAX == 6 ? 2 : 1;
  static constexpr int TN = 1;
  static constexpr int NT = 64 * (4 / TM) * (4 / TN);
  static constexpr int WPE = NT / 256;  // waves per SIMD
};

// One instantiation per largest product count of a launch (x1, x3, x6), so
// each is a kernel symbol of its own; planes may use fewer products.
// The tile count of a transform (its fringe workgroups follow).
__device__ inline int gemm_tiles(const GemmArgs& a) {
  const GemmCut cut = gemm_cut(a.RA, a.RB, a.K);
  return ((cut.am + BM - 1) / BM) * ((cut.bm + BN - 1) / BN) * a.nplanes;
}

// Transforms of one launch: G == 1 for a single transform, else up to
// MAX_GROUP (the batches of an ensemble), each transform's blocks mapped to
// tiles and XCDs as in its own launch.  Block lb of a transform is its tile
// lb (in the XCD-contiguous tile order) when lb < T, else its fringe
// workgroup lb - T; -1 marks the padding blocks of a merged launch.
template <int G>
struct GemmTable {
  BlockTable t;
  GemmArgs a[G];
};
template <int G>
__device__ inline int gemm_block(const GemmTable<G>& g, int& c) {
  if (G == 1) {
    c = 0;
    const int T = gemm_tiles(g.a[0]);
    const int b = blockIdx.x;
    return b < T ? remap_tiles(b, T) : b;
  }
  int l;
  if (!table_class(g.t, blockIdx.x, c, l)) return -1;  // padding between transforms
  const int T = gemm_tiles(g.a[c]);
  return l < T ? remap_tiles(l, T) : l;
}

template <int NPMAX, int G>
__global__ __launch_bounds__(KernelShape<NPMAX>::NT)
__attribute__((amdgpu_waves_per_eu(KernelShape<NPMAX>::WPE, KernelShape<NPMAX>::WPE)))
void dst_gemm_bf16_kernel(GemmTable<G> g) {
  constexpr int TM = KernelShape<NPMAX>::TM, TN = KernelShape<NPMAX>::TN, NT = KernelShape<NPMAX>::NT;
  extern __shared__ __attribute__((aligned(16))) uint16_t lds16[];
  int c;
  const int lb = gemm_block<G>(g, c);
  if (lb < 0) return;
  const GemmArgs& a = g.a[c];
  const GemmCut cut = gemm_cut(a.RA, a.RB, a.K);
  const int GA = (cut.am + BM - 1) / BM, GB = (cut.bm + BN - 1) / BN;
  const int T = GA * GB * a.nplanes;
  const bool fringe = lb >= T;
  int zi, ta = 0, tb = 0, part = 0;
  if (fringe) {
    zi = (lb - T) / cut.fpw;
    part = (lb - T) - zi * cut.fpw;
  } else {
    tile_of(lb, GA, GB, zi, ta, tb);
  }
  const long z = zi;
  const PlaneSel ps = select_plane(a, z);
  const bool lo = a.hp && a.hp[ps.m];
  const int np = lo ? a.np_lo : a.np_hi;
  const int opieces = lo ? a.op_lo : a.op_hi;
  if (fringe) {
    if (NPMAX == 1 || np == 1) gemm_fringe<false, 1, NT>(a, z, part, cut, ps, opieces, lds16);
    else if (NPMAX == 3 || np == 3) gemm_fringe<false, 2, NT>(a, z, part, cut, ps, opieces, lds16);
    else gemm_fringe<false, 3, NT>(a, z, part, cut, ps, opieces, lds16);
    return;
  }
  const int a0 = ta * BM;
  const int b0 = tb * BN;
  const bool x_in_a = !a.trans;
  const uint16_t* S = a.Sb + ps.s_off;

  __shared__ float lam_tile[256];
  const bool spectral = ps.sp.lamA && opieces != 0;
  const float lam_pre = spectral ? prefetch_spectral(a, ps.sp, a0, b0) : 0.0f;

  floatx16 acc[TM][TN];
  double sum[TM][TN][16];
  zero_acc<TM, TN>(acc, sum);
  // FP32 accumulation over K ~ 1e3 costs ~2e-7 relative, far below the
  // 3-product split error (~1e-5): only 6-product planes accumulate in FP64.
  if (NPMAX == 1 || np == 1) bf16_mainloop<1, false, TM, TN>(a, S, lds16, z, a0, b0, x_in_a, acc, sum);
  else if (NPMAX == 3 || np == 3) bf16_mainloop<3, false, TM, TN>(a, S, lds16, z, a0, b0, x_in_a, acc, sum);
  else if (a.acc64) bf16_mainloop<6, true, TM, TN>(a, S, lds16, z, a0, b0, x_in_a, acc, sum);
  else bf16_mainloop<6, false, TM, TN>(a, S, lds16, z, a0, b0, x_in_a, acc, sum);
  float* stage = reinterpret_cast<float*>(lds16);
  if (spectral && threadIdx.x < 256) lam_tile[threadIdx.x] = lam_pre;  // read after store_tile's first barrier
  const float* lam = spectral ? lam_tile : nullptr;
  if (NPMAX == 6 && np == 6 && a.acc64) store_tile<true, TM, TN>(a, z, a0, b0, acc, sum, opieces, stage, ps.sp, lam);
  else store_tile<false, TM, TN>(a, z, a0, b0, acc, sum, opieces, stage, ps.sp, lam);
}


// Pair-tile instance of the x1 / x3 transforms: tiles of 128 x 64, two
// workgroups per CU (see bf16_mainloop_pair); the fringe as in the other
// instances.
__device__ inline int gemm_tiles_pair(const GemmArgs& a) {
  const GemmCut cut = gemm_cut(a.RA, a.RB, a.K);
  return ((cut.am + BM - 1) / BM) * ((cut.bm + BNP - 1) / BNP) * a.nplanes;
}
template <int G>
__device__ inline int gemm_block_pair(const GemmTable<G>& g, int& c) {
  if (G == 1) {
    c = 0;
    const int T = gemm_tiles_pair(g.a[0]);
    const int b = blockIdx.x;
    return b < T ? remap_tiles(b, T) : b;
  }
  int l;
  if (!table_class(g.t, blockIdx.x, c, l)) return -1;  // padding between transforms
  const int T = gemm_tiles_pair(g.a[c]);
  return l < T ? remap_tiles(l, T) : l;
}

template <int NPMAX, int G>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4)))
void dst_gemm_bf16_pair_kernel(GemmTable<G> g) {
  static_assert(NPMAX == 1 || NPMAX == 3, "pair tiles serve the x1 and x3 transforms");
  extern __shared__ __attribute__((aligned(16))) uint16_t lds16[];
  int c;
  const int lb = gemm_block_pair<G>(g, c);
  if (lb < 0) return;
  const GemmArgs& a = g.a[c];
  const GemmCut cut = gemm_cut(a.RA, a.RB, a.K);
  const int GA = (cut.am + BM - 1) / BM, GB = (cut.bm + BNP - 1) / BNP;
  const int T = GA * GB * a.nplanes;
  const bool fringe = lb >= T;
  int zi, ta = 0, tb = 0, part = 0;
  if (fringe) {
    zi = (lb - T) / cut.fpw;
    part = (lb - T) - zi * cut.fpw;
  } else {
    tile_of(lb, GA, GB, zi, ta, tb);
  }
  const long z = zi;
  const PlaneSel ps = select_plane(a, z);
  const bool lo = a.hp && a.hp[ps.m];
  const int np = lo ? a.np_lo : a.np_hi;
  const int opieces = lo ? a.op_lo : a.op_hi;
  if (fringe) {
    if (NPMAX == 1 || np == 1) gemm_fringe<false, 1, 512, 4>(a, z, part, cut, ps, opieces, lds16);
    else gemm_fringe<false, 2, 512, 4>(a, z, part, cut, ps, opieces, lds16);
    return;
  }
  const int a0 = ta * BM;
  const int b0 = tb * BNP;
  const bool x_in_a = !a.trans;
  const uint16_t* S = a.Sb + ps.s_off;

  __shared__ float lam_tile[256];
  const bool spectral = ps.sp.lamA && opieces != 0;
  const float lam_pre = spectral ? prefetch_spectral(a, ps.sp, a0, b0) : 0.0f;
  floatx16 acc[1][1];
  double sum[1][1][16];
  zero_acc<1, 1>(acc, sum);
  if (NPMAX == 1 || np == 1) bf16_mainloop_pair<1>(a, S, lds16, z, a0, b0, x_in_a, acc);
  else bf16_mainloop_pair<3>(a, S, lds16, z, a0, b0, x_in_a, acc);
  float* stage = reinterpret_cast<float*>(lds16);
  if (spectral && threadIdx.x < 256) lam_tile[threadIdx.x] = lam_pre;  // read after store_tile's first barrier
  const float* lam = spectral ? lam_tile : nullptr;
  store_tile<false, 1, 1, 2, BNP>(a, z, a0, b0, acc, sum, opieces, stage, ps.sp, lam);
}

// ---------------------------------------------------------------------------
// FP32 kernel (plain step): register-staged double buffer, FP64 flush every
// 4 k-tiles of 16.
// ---------------------------------------------------------------------------
template <int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void dst_gemm_f32_kernel(GemmTable<G> g) {
  __shared__ __attribute__((aligned(16))) float lds[2 * 2 * TILE32];
  int c;
  const int lb = gemm_block<G>(g, c);
  if (lb < 0) return;
  const GemmArgs& a = g.a[c];
  const GemmCut cut = gemm_cut(a.RA, a.RB, a.K);
  const int GA = (cut.am + BM - 1) / BM, GB = (cut.bm + BN - 1) / BN;
  const int T = GA * GB * a.nplanes;
  if (lb >= T) {
    const int zf = (lb - T) / cut.fpw;
    gemm_fringe<true, 1, 256>(a, zf, (lb - T) - zf * cut.fpw, cut, select_plane(a, zf), 0, lds);
    return;
  }
  int zi, ta, tb;
  tile_of(lb, GA, GB, zi, ta, tb);
  const long z = zi;
  const PlaneSel ps = select_plane(a, z);
  const int a0 = ta * BM;
  const int b0 = tb * BN;
  const bool x_in_a = !a.trans;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const float* X = a.Xf + z * a.x_batch;
  const float* S = a.Sf + ps.s_off;
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
  double sum[2][2][16];
  zero_acc<2, 2>(acc, sum);
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
  // one k-tile: prefetch the next into registers, MFMA the current from LDS
  // into `w`, then park the prefetch in the other LDS buffer
  auto ktile = [&](int kt, floatx16(&w)[2][2]) {
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
          w[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i], fb[j], w[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) {
      float* tb = lds + (buf ^ 1) * 2 * TILE32;
      *reinterpret_cast<float4*>(tb + xo + st0) = rx0;
      *reinterpret_cast<float4*>(tb + xo + st1) = rx1;
      *reinterpret_cast<float4*>(tb + so + st0) = rs0;
      *reinterpret_cast<float4*>(tb + so + st1) = rs1;
    }
    __syncthreads();
  };
  if (!acc64) {
    for (int kt = 0; kt < nk; ++kt) ktile(kt, acc);
  } else {
    // FP64 running sum over windows of 4 k-tiles, each window's MFMA chain
    // starting from zero (no FP32 accumulator carried around the loop)
    int kt = 0;
    for (; kt + 4 <= nk; kt += 4) {
      floatx16 w[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) w[i][j] = floatx16{};
#pragma unroll
      for (int f = 0; f < 4; ++f) ktile(kt + f, w);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) sum[i][j][r] += (double)w[i][j][r];
    }
    for (; kt < nk; ++kt) ktile(kt, acc);
  }
  if (acc64) store_tile<true, 2, 2>(a, z, a0, b0, acc, sum, 0, lds, ps.sp);
  else store_tile<false, 2, 2>(a, z, a0, b0, acc, sum, 0, lds, ps.sp);
}


template __global__ void dst_gemm_bf16_kernel<1, 1>(GemmTable<1>);
template __global__ void dst_gemm_bf16_kernel<3, 1>(GemmTable<1>);
template __global__ void dst_gemm_bf16_kernel<6, 1>(GemmTable<1>);
template __global__ void dst_gemm_bf16_kernel<1, MAX_GROUP>(GemmTable<MAX_GROUP>);
template __global__ void dst_gemm_bf16_kernel<3, MAX_GROUP>(GemmTable<MAX_GROUP>);
template __global__ void dst_gemm_bf16_kernel<6, MAX_GROUP>(GemmTable<MAX_GROUP>);
template __global__ void dst_gemm_bf16_pair_kernel<1, 1>(GemmTable<1>);
template __global__ void dst_gemm_bf16_pair_kernel<3, 1>(GemmTable<1>);
template __global__ void dst_gemm_bf16_pair_kernel<1, MAX_GROUP>(GemmTable<MAX_GROUP>);
template __global__ void dst_gemm_bf16_pair_kernel<3, MAX_GROUP>(GemmTable<MAX_GROUP>);

hipError_t prepare_dst_gemm() {
  // up to 144 KiB of the 160 KiB LDS of a gfx950 CU (3 pieces, 3 stages);
  // set outside any stream capture, before the first launch
  static hipError_t status = [] {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_kernel<1, 1>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_lds_bytes(1));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_kernel<3, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_lds_bytes(3));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_kernel<6, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_lds_bytes(6));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_kernel<1, MAX_GROUP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_lds_bytes(1));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_kernel<3, MAX_GROUP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_lds_bytes(3));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_kernel<6, MAX_GROUP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_lds_bytes(6));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_pair_kernel<1, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_pair_lds_bytes(1));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_pair_kernel<3, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_pair_lds_bytes(3));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_pair_kernel<1, MAX_GROUP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_pair_lds_bytes(1));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(dst_gemm_bf16_pair_kernel<3, MAX_GROUP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dst_gemm_pair_lds_bytes(3));
    return e;
  }();
  return status;
}

namespace {
// Blocks of one transform: tiles of 128 x 128 (or 128 x 64 with pair tiles)
// and the fringe workgroups.
int gemm_blocks(const GemmArgs& a, bool pair = false) {
  const GemmCut cut = gemm_cut(a.RA, a.RB, a.K);
  const int bn = pair ? BNP : BN;
  const int tiles = ((cut.bm + bn - 1) / bn) * ((cut.am + BM - 1) / BM) * a.nplanes;
  return tiles + cut.fpw * a.nplanes;
}
int max_products(const GemmArgs& a) { return a.np_lo > a.np_hi ? a.np_lo : a.np_hi; }
// Compute units of the current device (looked up once per device; several
// host threads may race to fill an entry, and they store the same value).
int cu_count() {
  static std::atomic<int> cus[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = cus[dev].load(std::memory_order_relaxed);
  if (!c) {
    int n = 0;
    c = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0 ? n : 256;
    cus[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}
// The x1 / x3 transforms run on pair tiles when their 128 x 128 workgroups
// (tiles and fringe, `blocks`) would take more than one round of the CUs:
// measured, the 256-member ensemble's step 3.41 -> 3.23 ms, while the 1024^2
// model's transforms, one round of 252 workgroups, are faster whole
// (22.2 against 25.5 us for a x3 launch).
bool use_pair(const GemmArgs& a, int max_np, long blocks) { return !a.Xf && max_np <= 3 && blocks > cu_count(); }
}  // namespace

void launch_dst_gemm_group(const GemmArgs* args, const int* batch, int n, hipStream_t st) {
  GemmTable<MAX_GROUP> g{};
  int max_np = 0;
  for (int i = 0; i < n; ++i) {
    const int np = max_products(args[i]);
    max_np = np > max_np ? np : max_np;
  }
  long whole = 0;  // workgroups with 128 x 128 tiles
  for (int i = 0; i < n; ++i) {
    GemmArgs t = args[i];
    t.nplanes = batch[i];
    whole += gemm_blocks(t);
  }
  const bool pair = use_pair(args[0], max_np, whole);
  for (int i = 0; i < n; ++i) {
    g.a[i] = args[i];
    g.a[i].nplanes = batch[i];
    table_add(g.t, gemm_blocks(g.a[i], pair), 1, 1);
  }
  const dim3 grid(g.t.start[g.t.n]);
  if (args[0].Xf) {
    hipLaunchKernelGGL(dst_gemm_f32_kernel<MAX_GROUP>, grid, dim3(256), 0, st, g);
    return;
  }
  if (pair) {
    const size_t lp = dst_gemm_pair_lds_bytes(max_np);
    if (max_np == 1) hipLaunchKernelGGL((dst_gemm_bf16_pair_kernel<1, MAX_GROUP>), grid, dim3(512), lp, st, g);
    else hipLaunchKernelGGL((dst_gemm_bf16_pair_kernel<3, MAX_GROUP>), grid, dim3(512), lp, st, g);
    return;
  }
  const size_t lds = dst_gemm_lds_bytes(max_np);
  if (max_np == 1)
    hipLaunchKernelGGL((dst_gemm_bf16_kernel<1, MAX_GROUP>), grid, dim3(KernelShape<1>::NT), lds, st, g);
  else if (max_np == 3)
    hipLaunchKernelGGL((dst_gemm_bf16_kernel<3, MAX_GROUP>), grid, dim3(KernelShape<3>::NT), lds, st, g);
  else
    hipLaunchKernelGGL((dst_gemm_bf16_kernel<6, MAX_GROUP>), grid, dim3(KernelShape<6>::NT), lds, st, g);
}

void launch_dst_gemm(const GemmArgs& args, int batch, hipStream_t st) {
  GemmTable<1> g{};
  g.a[0] = args;
  g.a[0].nplanes = batch;
  const GemmArgs& a = g.a[0];
  const int max_np = max_products(a);
  const bool pair = use_pair(a, max_np, gemm_blocks(a));
  const int blocks = gemm_blocks(a, pair);
  if (a.Xf) {
    hipLaunchKernelGGL(dst_gemm_f32_kernel<1>, dim3(blocks), dim3(256), 0, st, g);
    return;
  }
  if (pair) {
    const size_t lp = dst_gemm_pair_lds_bytes(max_np);
    if (max_np == 1) hipLaunchKernelGGL((dst_gemm_bf16_pair_kernel<1, 1>), dim3(blocks), dim3(512), lp, st, g);
    else hipLaunchKernelGGL((dst_gemm_bf16_pair_kernel<3, 1>), dim3(blocks), dim3(512), lp, st, g);
    return;
  }
  const size_t lds = dst_gemm_lds_bytes(max_np);
  if (max_np == 1)
    hipLaunchKernelGGL((dst_gemm_bf16_kernel<1, 1>), dim3(blocks), dim3(KernelShape<1>::NT), lds, st, g);
  else if (max_np == 3)
    hipLaunchKernelGGL((dst_gemm_bf16_kernel<3, 1>), dim3(blocks), dim3(KernelShape<3>::NT), lds, st, g);
  else
    hipLaunchKernelGGL((dst_gemm_bf16_kernel<6, 1>), dim3(blocks), dim3(KernelShape<6>::NT), lds, st, g);
}

}  // namespace somax
