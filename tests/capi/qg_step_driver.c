/* A plain-C host of the somax HIP library: no Python, no torch.  It reads a
 * QG description and an initial PV field from a binary file, builds a plan
 * through the C ABI of include/somaxhip.h, advances nsteps SSP-RK3 steps on
 * the device and writes the result.  tests/test_gpu_capi.py compiles it with
 * gcc and compares its output with the Python path's, which runs the same
 * kernels.
 *
 * Input (little-endian): int32 nx, ny, nl, nsteps, precision; float64 dx, dy,
 * dt, nu, nu4, kappa, tau0; float64 cl2m[nl*nl], cm2l[nl*nl], lambdas[nl],
 * beta_y_row[ny+2], wind_row[ny+2], wind_scale[nl]; q[nl*(ny+2)*(nx+2)], float32
 * or, for precision SOMAXHIP_PREC_FP64, float64 (stepped by somaxhip_qg_step_f64).
 * Output: q after nsteps steps, in the input's type. */
#include <stdio.h>
#include <stdlib.h>

#include <hip/hip_runtime_api.h>

#include "somaxhip.h"

static int fail(const char* what, int rc) {
  char msg[512];
  somaxhip_last_error(msg, sizeof msg);
  fprintf(stderr, "%s failed (%d): %s\n", what, rc, msg);
  return 1;
}

static void* read_n(FILE* f, size_t bytes) {
  void* p = malloc(bytes ? bytes : 1);
  if (!p || fread(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "short input\n");
    exit(2);
  }
  return p;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int* hdr = (int*)read_n(f, 5 * sizeof(int));
  const int nx = hdr[0], ny = hdr[1], nl = hdr[2], nsteps = hdr[3], precision = hdr[4];
  double* sc = (double*)read_n(f, 7 * sizeof(double));
  double* cl2m = (double*)read_n(f, (size_t)nl * nl * sizeof(double));
  double* cm2l = (double*)read_n(f, (size_t)nl * nl * sizeof(double));
  double* lambdas = (double*)read_n(f, (size_t)nl * sizeof(double));
  double* beta = (double*)read_n(f, (size_t)(ny + 2) * sizeof(double));
  double* wind = (double*)read_n(f, (size_t)(ny + 2) * sizeof(double));
  double* wscale = (double*)read_n(f, (size_t)nl * sizeof(double));
  const size_t n = (size_t)nl * (ny + 2) * (nx + 2);
  const int f64 = precision == SOMAXHIP_PREC_FP64;
  const size_t esz = f64 ? sizeof(double) : sizeof(float);
  void* q = read_n(f, n * esz);
  fclose(f);

  int rc = somaxhip_init();
  if (rc) return fail("somaxhip_init", rc);
  somaxhip_qg_desc d = {nx, ny, nl, sc[0], sc[1], cl2m, cm2l, lambdas, beta, wind, wscale,
                        /*zero_psi_ghost*/ 1, SOMAXHIP_ADV_ARAKAWA, precision, /*mass_correction*/ 0, NULL, NULL};
  somaxhip_qg_plan plan;
  if ((rc = somaxhip_qg_plan_create(&d, &plan))) return fail("somaxhip_qg_plan_create", rc);
  const size_t wsb = somaxhip_qg_workspace_bytes(plan);
  void* dq = NULL;
  void* ws = NULL;
  if (hipMalloc(&dq, n * esz) != hipSuccess || hipMalloc(&ws, wsb) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemcpy(dq, q, n * esz, hipMemcpyHostToDevice) != hipSuccess) return 1;
  somaxhip_qg_params prm = {sc[3], sc[4], sc[5], sc[6]};
  if (f64) {
    if ((rc = somaxhip_qg_step_f64(plan, (double*)dq, NULL, nsteps, sc[2], &prm, ws, NULL)))
      return fail("somaxhip_qg_step_f64", rc);
  } else if ((rc = somaxhip_qg_step(plan, (float*)dq, NULL, nsteps, sc[2], &prm, ws, NULL))) {
    return fail("somaxhip_qg_step", rc);
  }
  if (hipMemcpy(q, dq, n * esz, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  somaxhip_qg_plan_destroy(plan);
  hipFree(dq);
  hipFree(ws);
  FILE* o = fopen(argv[2], "wb");
  if (!o || fwrite(q, esz, n, o) != n) return 2;
  fclose(o);
  printf("stepped %d x %d x %d for %d steps\n", nl, ny + 2, nx + 2, nsteps);
  return 0;
}
