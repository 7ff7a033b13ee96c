// Capacitance-matrix solve of the Helmholtz problem on a masked domain.
//
// The rectangular DST-I inversion of a QG plan (identity modal matrices, one
// Helmholtz parameter per plane, ghost ring kept) is G.  With B the points
// outside the mask that touch it and C_p = (G_p on B x B)^-1 per plane, a
// solve is
//   r0 = r * mask;  psi1 = G r0;  r0[B] = -C psi1[B];  psi = (G r0) * mask,
// so psi vanishes on B and solves (lap - lam_p) psi = r on the mask: two
// rectangular inversions plus an FP64 matrix-vector product per plane.

#include <algorithm>
#include <new>
#include <string>

#include "common.h"
#include "kernels.h"
#include "somaxhip.h"

namespace somax {
namespace {

constexpr int CMM_THREADS = 256;

// out[i] = in[i] * mask[i mod plane] over planes of `plane` values (in place
// when in == out).
__global__ void __launch_bounds__(CMM_THREADS) cmm_mask_kernel(const float* in, float* out,
                                                               const float* __restrict__ mask, long plane,
                                                               long total) {
  const long i = (long)blockIdx.x * CMM_THREADS + threadIdx.x;
  if (i >= total) return;
  out[i] = in[i] * mask[i % plane];
}

// One wave per boundary row i of plane p: r0[p][b_i] = -sum_j C_p[i][j] psi1[p][b_j]
// in FP64, lanes striding over j, then a fixed butterfly.  grid (ceil(nb/4), P).
__global__ void __launch_bounds__(CMM_THREADS) cmm_alpha_kernel(const float* __restrict__ psi1,
                                                                float* __restrict__ r0,
                                                                const double* __restrict__ cap,
                                                                const int* __restrict__ bidx, int nb, long plane) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (CMM_THREADS / 64) + (threadIdx.x >> 6);
  const int p = blockIdx.y;
  if (i >= nb) return;  // whole waves leave together; no block-level barrier follows
  const double* row = cap + ((size_t)p * nb + i) * nb;
  const float* src = psi1 + (size_t)p * plane;
  double acc = 0.0;
  for (int j = lane; j < nb; j += 64) acc += row[j] * (double)src[bidx[j]];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) r0[(size_t)p * plane + bidx[i]] = (float)(-acc);
}

}  // namespace
}  // namespace somax

using namespace somax;

struct somaxhip_cmm_s {
  somaxhip_qg_plan plan = nullptr;
  int P = 0, Ny = 0, Nx = 0, nb = 0;
  long plane = 0;
  size_t plan_ws = 0;  // bytes of the plan's workspace, rounded up to 256
  void* mem = nullptr;
  float* mask = nullptr;
  int* bidx = nullptr;
  double* cap = nullptr;
  ~somaxhip_cmm_s() {
    if (mem) (void)hipFree(mem);
  }
};

extern "C" {

int somaxhip_cmm_create(somaxhip_qg_plan plan, const float* mask, int nb, const int* bidx, const double* cap,
                        somaxhip_cmm* out) {
  if (!plan || !mask || !out || nb < 0 || (nb > 0 && (!bidx || !cap))) {
    set_error("somaxhip_cmm_create: null plan, mask, boundary table or output, or negative boundary size");
    return SOMAXHIP_ERR_ARG;
  }
  *out = nullptr;
  auto c = new (std::nothrow) somaxhip_cmm_s();
  if (!c) {
    set_error("somaxhip_cmm_create: out of host memory");
    return SOMAXHIP_ERR_ALLOC;
  }
  c->plan = plan;
  if (qg_plan_shape(plan, &c->P, &c->Ny, &c->Nx) != SOMAXHIP_OK) {
    set_error("somaxhip_cmm_create: the QG plan computes in FP64; the masked solve takes an FP32 / BF16 plan");
    delete c;
    return SOMAXHIP_ERR_ARG;
  }
  c->plane = (long)c->Ny * c->Nx;
  c->nb = nb;
  for (int i = 0; i < nb; ++i) {
    if (bidx[i] < 0 || bidx[i] >= c->plane || mask[bidx[i]] != 0.0f) {
      set_error("somaxhip_cmm_create: boundary point " + std::to_string(i) +
                " is outside the field or inside the mask");
      delete c;
      return SOMAXHIP_ERR_ARG;
    }
  }
  c->plan_ws = (somaxhip_qg_workspace_bytes(plan) + 255) & ~(size_t)255;
  const size_t mbytes = ((size_t)c->plane * sizeof(float) + 255) & ~(size_t)255;
  const size_t bbytes = ((size_t)std::max(nb, 1) * sizeof(int) + 255) & ~(size_t)255;
  const size_t cbytes = (size_t)c->P * nb * nb * sizeof(double);
  if (hipMalloc(&c->mem, mbytes + bbytes + std::max(cbytes, (size_t)256)) != hipSuccess) {
    set_error("somaxhip_cmm_create: device allocation failed");
    delete c;
    return SOMAXHIP_ERR_ALLOC;
  }
  char* m = static_cast<char*>(c->mem);
  c->mask = reinterpret_cast<float*>(m);
  c->bidx = reinterpret_cast<int*>(m + mbytes);
  c->cap = reinterpret_cast<double*>(m + mbytes + bbytes);
  hipError_t e = hipMemcpy(c->mask, mask, c->plane * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && nb) e = hipMemcpy(c->bidx, bidx, nb * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && nb) e = hipMemcpy(c->cap, cap, cbytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("somaxhip_cmm_create: ") + hipGetErrorString(e));
    delete c;
    return SOMAXHIP_ERR_HIP;
  }
  *out = c;
  return SOMAXHIP_OK;
}

int somaxhip_cmm_destroy(somaxhip_cmm c) {
  delete c;
  return SOMAXHIP_OK;
}

size_t somaxhip_cmm_workspace_bytes(somaxhip_cmm c) {
  if (!c) return 0;
  return c->plan_ws + 2 * (((size_t)c->P * c->plane * sizeof(float) + 255) & ~(size_t)255);
}

int somaxhip_cmm_solve(somaxhip_cmm c, const float* rhs, float* psi, void* ws, somaxhip_stream_t stream) {
  if (!c || !rhs || !psi || !ws) {
    set_error("somaxhip_cmm_solve: null solver, field or workspace");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const long total = (long)c->P * c->plane;
  const size_t fbytes = ((size_t)total * sizeof(float) + 255) & ~(size_t)255;
  float* r0 = reinterpret_cast<float*>(static_cast<char*>(ws) + c->plan_ws);
  float* psi1 = reinterpret_cast<float*>(static_cast<char*>(ws) + c->plan_ws + fbytes);
  const unsigned blocks = (unsigned)((total + CMM_THREADS - 1) / CMM_THREADS);
  hipLaunchKernelGGL(cmm_mask_kernel, dim3(blocks), dim3(CMM_THREADS), 0, st, rhs, r0, (const float*)c->mask,
                     c->plane, total);
  int rc = somaxhip_qg_invert(c->plan, r0, psi1, 0, ws, stream);
  if (rc) return rc;
  if (c->nb) {
    const dim3 g((unsigned)((c->nb + CMM_THREADS / 64 - 1) / (CMM_THREADS / 64)), (unsigned)c->P);
    hipLaunchKernelGGL(cmm_alpha_kernel, g, dim3(CMM_THREADS), 0, st, (const float*)psi1, r0,
                       (const double*)c->cap, (const int*)c->bidx, c->nb, c->plane);
  }
  rc = somaxhip_qg_invert(c->plan, r0, psi, 0, ws, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(cmm_mask_kernel, dim3(blocks), dim3(CMM_THREADS), 0, st, (const float*)psi, psi,
                     (const float*)c->mask, c->plane, total);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

}  // extern "C"
