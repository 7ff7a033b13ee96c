// Field statistics for the between-chunk diagnostics of a run: per float32
// field the minimum and maximum over its non-NaN values, their sum (FP64),
// and the counts of NaN and of non-finite (NaN or +-Inf) values.  Two
// launches reduce in a fixed order (per-block partials, then one block per
// field), so results are deterministic for a given field size and only five
// doubles per field travel to the host.

#include <algorithm>

#include "common.h"
#include "somaxhip.h"

namespace somax {
namespace {

constexpr int ST_THREADS = 256;
constexpr int ST_WAVES = ST_THREADS / 64;
constexpr int ST_MAX_BLOCKS = 1024;          // partial slots per field
constexpr long ST_PER_BLOCK = 4096;          // elements per block before the cap
constexpr int ST_OUT = 5;                    // min, max, sum, NaN count, non-finite count

struct StatFields {
  const float* p[SOMAXHIP_STATS_MAX_FIELDS];
  unsigned long long n[SOMAXHIP_STATS_MAX_FIELDS];
};

struct Acc {
  float mn, mx;
  double sum;
  unsigned long long nnan, nbad;
};

__host__ __device__ inline int stat_blocks(unsigned long long n) {
  const unsigned long long b = (n + ST_PER_BLOCK - 1) / ST_PER_BLOCK;
  return b < 1 ? 1 : (b > (unsigned long long)ST_MAX_BLOCKS ? ST_MAX_BLOCKS : (int)b);
}

__device__ inline Acc acc_identity() { return Acc{__builtin_inff(), -__builtin_inff(), 0.0, 0ull, 0ull}; }

// Classification by the IEEE-754 bits, independent of floating-point
// contraction or fast-math settings.
__device__ inline void acc_add(Acc& a, float x) {
  const uint32_t b = __float_as_uint(x);
  const bool nonfinite = (b & 0x7f800000u) == 0x7f800000u;
  const bool nan = nonfinite && (b & 0x007fffffu) != 0u;
  a.nbad += nonfinite ? 1ull : 0ull;
  a.nnan += nan ? 1ull : 0ull;
  if (!nan) {
    a.mn = fminf(a.mn, x);
    a.mx = fmaxf(a.mx, x);
    a.sum += (double)x;
  }
}

__device__ inline void acc_merge(Acc& a, const Acc& b) {
  a.mn = fminf(a.mn, b.mn);
  a.mx = fmaxf(a.mx, b.mx);
  a.sum += b.sum;
  a.nnan += b.nnan;
  a.nbad += b.nbad;
}

// Butterfly over the 64 lanes of a wave: every lane ends with the same,
// order-fixed result.
__device__ inline void acc_wave(Acc& a) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    Acc o;
    o.mn = __shfl_xor(a.mn, off, 64);
    o.mx = __shfl_xor(a.mx, off, 64);
    o.sum = __shfl_xor(a.sum, off, 64);
    o.nnan = __shfl_xor(a.nnan, off, 64);
    o.nbad = __shfl_xor(a.nbad, off, 64);
    acc_merge(a, o);
  }
}

// Reduces one Acc per thread of a 256-thread block; thread 0 returns the total.
__device__ inline Acc acc_block(Acc a) {
  __shared__ Acc part[ST_WAVES];
  acc_wave(a);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) part[wave] = a;
  __syncthreads();
  Acc t = part[0];
  if (threadIdx.x == 0)
    for (int w = 1; w < ST_WAVES; ++w) acc_merge(t, part[w]);
  return t;
}

// grid (ST_MAX_BLOCKS or fewer, nfields): block b of field f covers the
// float4 vectors b * 256 + t strided by the field's block count x 256.
__global__ void __launch_bounds__(ST_THREADS) stats_partial_kernel(StatFields fs, double* __restrict__ partial) {
  const int f = blockIdx.y;
  const unsigned long long n = fs.n[f];
  const int nb = stat_blocks(n);
  if ((int)blockIdx.x >= nb) return;
  const float* __restrict__ x = fs.p[f];
  Acc a = acc_identity();
  const unsigned long long stride = (unsigned long long)nb * ST_THREADS;
  const unsigned long long t0 = (unsigned long long)blockIdx.x * ST_THREADS + threadIdx.x;
  if ((reinterpret_cast<uintptr_t>(x) & 15u) == 0u) {
    const unsigned long long nv = n / 4;
    const float4* __restrict__ x4 = reinterpret_cast<const float4*>(x);
    unsigned long long i = t0;
    // four independent 16-byte loads in flight per lane, then their values
    for (; i + 3 * stride < nv; i += 4 * stride) {
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = x4[i + u * stride];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        acc_add(a, v[u].x);
        acc_add(a, v[u].y);
        acc_add(a, v[u].z);
        acc_add(a, v[u].w);
      }
    }
    for (; i < nv; i += stride) {
      const float4 v = x4[i];
      acc_add(a, v.x);
      acc_add(a, v.y);
      acc_add(a, v.z);
      acc_add(a, v.w);
    }
    const unsigned long long tail = nv * 4 + t0;
    if (tail < n) acc_add(a, x[tail]);  // t0 < 4 only, since n - 4 nv < 4
  } else {
    for (unsigned long long i = t0; i < n; i += stride) acc_add(a, x[i]);
  }
  a = acc_block(a);
  if (threadIdx.x == 0) {
    double* o = partial + ((size_t)f * ST_MAX_BLOCKS + blockIdx.x) * ST_OUT;
    o[0] = a.mn;
    o[1] = a.mx;
    o[2] = a.sum;
    o[3] = (double)a.nnan;
    o[4] = (double)a.nbad;
  }
}

// grid (nfields): one block folds the field's partials in index order.
__global__ void __launch_bounds__(ST_THREADS) stats_final_kernel(StatFields fs, const double* __restrict__ partial,
                                                                  double* __restrict__ out) {
  const int f = blockIdx.x;
  const int nb = stat_blocks(fs.n[f]);
  Acc a = acc_identity();
  for (int b = threadIdx.x; b < nb; b += ST_THREADS) {
    const double* p = partial + ((size_t)f * ST_MAX_BLOCKS + b) * ST_OUT;
    Acc o{(float)p[0], (float)p[1], p[2], (unsigned long long)p[3], (unsigned long long)p[4]};
    acc_merge(a, o);
  }
  a = acc_block(a);
  if (threadIdx.x == 0) {
    double* o = out + (size_t)f * ST_OUT;
    o[0] = a.mn;
    o[1] = a.mx;
    o[2] = a.sum;
    o[3] = (double)a.nnan;
    o[4] = (double)a.nbad;
  }
}

}  // namespace
}  // namespace somax

using namespace somax;

extern "C" {

size_t somaxhip_stats_workspace_bytes(int nfields) {
  if (nfields < 1 || nfields > SOMAXHIP_STATS_MAX_FIELDS) return 0;
  return (size_t)nfields * ST_MAX_BLOCKS * ST_OUT * sizeof(double);
}

int somaxhip_field_stats(const float* const* fields, const size_t* counts, int nfields, double* stats, void* ws,
                         somaxhip_stream_t stream) {
  if (nfields < 1 || nfields > SOMAXHIP_STATS_MAX_FIELDS) {
    set_error("somaxhip_field_stats: nfields must be in [1, " + std::to_string(SOMAXHIP_STATS_MAX_FIELDS) + "]");
    return SOMAXHIP_ERR_ARG;
  }
  if (!fields || !counts || !stats || !ws) {
    set_error("somaxhip_field_stats: null field list, counts, output or workspace");
    return SOMAXHIP_ERR_ARG;
  }
  StatFields fs{};
  for (int f = 0; f < nfields; ++f) {
    if (!fields[f] && counts[f] > 0) {
      set_error("somaxhip_field_stats: null field " + std::to_string(f));
      return SOMAXHIP_ERR_ARG;
    }
    fs.p[f] = fields[f];
    fs.n[f] = counts[f];
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int nbmax = 1;
  for (int f = 0; f < nfields; ++f) nbmax = std::max(nbmax, stat_blocks(fs.n[f]));
  double* partial = static_cast<double*>(ws);
  hipLaunchKernelGGL(stats_partial_kernel, dim3(nbmax, nfields), dim3(ST_THREADS), 0, st, fs, partial);
  hipLaunchKernelGGL(stats_final_kernel, dim3(nfields), dim3(ST_THREADS), 0, st, fs, (const double*)partial, stats);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

}  // extern "C"
