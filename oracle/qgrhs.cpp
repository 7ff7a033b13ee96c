/*
 * Restatement in C++ (OpenMP over rows) of the stencil part of the QG
 * right-hand side of oracle/qg.py:rhs: in FP64 for the large-grid parity
 * checks, in FP32 for the CPU baseline's step (oracle/cpu.py:FftCpuStep).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by oracle/qg.py (rhs(..., fast=True)),
 * through it by tests, and by bench.py's CPU leg; the product never loads it.
 * It follows somax/_src/models/qg/baroclinic.py:161-190 (vector_field) after
 * the inversion: the same operators as oracle/stencils.py -- the Arakawa
 * (1966) Jacobian, the flux-form WENO-Z advection with corner-streamfunction
 * face velocities, the 5-point Laplacian and the biharmonic -- evaluated term
 * by term in the same order and with the same operand order as the NumPy
 * code, and compiled without floating-point contraction, so that the FP64
 * instance gives NumPy's bits (checked by tests/test_oracle_cpu.py).
 *
 *   dq = adv + tau0 wind_scale[l] wind                  (adv = -J(psi, qt) or
 *        - kappa lap(psi[nl-1])   (last layer)           -div(u qt), interior)
 *        + nu lap(q) - nu4 lap(lap(q))
 *
 * qt = q + beta_y is passed in; q is the state with its boundary condition
 * applied; psi_wall[l] is the wall value of the corner streamfunction.
 */
#include <omp.h>

#include <cmath>
#include <cstddef>
#include <vector>

namespace {

template <class T>
T weno5z(T qm2, T qm1, T q0, T qp1, T qp2) {
  const T eps = (T)1e-20;  // oracle/stencils.py:WENO_EPS
  const T p0 = ((T)2 * qm2 - (T)7 * qm1 + (T)11 * q0) / (T)6;
  const T p1 = (-qm1 + (T)5 * q0 + (T)2 * qp1) / (T)6;
  const T p2 = ((T)2 * q0 + (T)5 * qp1 - qp2) / (T)6;
  const T d0 = qm2 - (T)2 * qm1 + q0, e0 = qm2 - (T)4 * qm1 + (T)3 * q0;
  const T d1 = qm1 - (T)2 * q0 + qp1, e1 = qm1 - qp1;
  const T d2 = q0 - (T)2 * qp1 + qp2, e2 = (T)3 * q0 - (T)4 * qp1 + qp2;
  const T c13 = (T)(13.0 / 12.0);
  const T b0 = c13 * (d0 * d0) + (T)0.25 * (e0 * e0);
  const T b1 = c13 * (d1 * d1) + (T)0.25 * (e1 * e1);
  const T b2 = c13 * (d2 * d2) + (T)0.25 * (e2 * e2);
  const T tau = std::fabs(b0 - b2);
  const T a0 = (T)0.1 * ((T)1 + tau / (b0 + eps));
  const T a1 = (T)0.6 * ((T)1 + tau / (b1 + eps));
  const T a2 = (T)0.3 * ((T)1 + tau / (b2 + eps));
  return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
}

template <class T>
T weno3z(T qm1, T q0, T qp1) {
  const T eps = (T)1e-20;
  const T p0 = (-qm1 + (T)3 * q0) / (T)2;
  const T p1 = (q0 + qp1) / (T)2;
  const T b0 = (q0 - qm1) * (q0 - qm1);
  const T b1 = (qp1 - q0) * (qp1 - q0);
  const T tau = std::fabs(b0 - b1);
  const T a0 = (T)(1.0 / 3.0) * ((T)1 + tau / (b0 + eps));
  const T a1 = (T)(2.0 / 3.0) * ((T)1 + tau / (b1 + eps));
  return (a0 * p0 + a1 * p1) / (a0 + a1);
}

// Upwind WENO value on face f+1/2 (f = 1..n-1) of a line of n interior cells
// c[k*s], k = 1..n (c points one stride before cell 1), with velocity vel:
// WENO5 where the five upwind cells are interior, else WENO3 where three are,
// else first-order upwind (oracle/stencils.py:_face_values_1d).
template <class T>
T face_value(const T* c, long s, int n, int f, T vel) {
  auto C = [&](int k) { return c[(long)k * s]; };
  if (vel > (T)0) {
    if (f - 2 >= 1 && f + 2 <= n) return weno5z(C(f - 2), C(f - 1), C(f), C(f + 1), C(f + 2));
    if (f - 1 >= 1 && f + 1 <= n) return weno3z(C(f - 1), C(f), C(f + 1));
    return C(f);
  }
  const int g = f + 1;
  if (g - 2 >= 1 && g + 2 <= n) return weno5z(C(f + 3), C(f + 2), C(f + 1), C(f), C(f - 1));
  if (g - 1 >= 1 && g + 1 <= n) return weno3z(C(f + 2), C(f + 1), C(f));
  return C(f + 1);
}

// Laplacian of one plane on the interior, ghost ring 0 (stencils.laplacian).
template <class T>
void laplacian(const T* f, T* out, int Ny, int Nx, T dx2, T dy2) {
#pragma omp parallel for schedule(static)
  for (long j = 0; j < Ny; ++j)
    for (long i = 0; i < Nx; ++i) {
      T v = 0;
      if (j > 0 && j < Ny - 1 && i > 0 && i < Nx - 1) {
        const T c = f[j * Nx + i];
        v = (f[j * Nx + i + 1] - (T)2 * c + f[j * Nx + i - 1]) / dx2 +
            (f[(j + 1) * Nx + i] - (T)2 * c + f[(j - 1) * Nx + i]) / dy2;
      }
      out[j * Nx + i] = v;
    }
}

// -J(p, q) on the interior, ghost ring 0 (stencils.arakawa_jacobian).
template <class T>
void arakawa(const T* p, const T* q, T* out, int Ny, int Nx, T dx, T dy) {
  const T den = (T)12 * dx * dy;
#pragma omp parallel for schedule(static)
  for (long j = 0; j < Ny; ++j)
    for (long i = 0; i < Nx; ++i) {
      if (j == 0 || j == Ny - 1 || i == 0 || i == Nx - 1) {
        out[j * Nx + i] = 0;
        continue;
      }
      auto P = [&](int dj, int di) { return p[(j + dj) * Nx + i + di]; };
      auto Q = [&](int dj, int di) { return q[(j + dj) * Nx + i + di]; };
      const T pE = P(0, 1), pW = P(0, -1), pN = P(1, 0), pS = P(-1, 0);
      const T pNE = P(1, 1), pNW = P(1, -1), pSE = P(-1, 1), pSW = P(-1, -1);
      const T qE = Q(0, 1), qW = Q(0, -1), qN = Q(1, 0), qS = Q(-1, 0);
      const T qNE = Q(1, 1), qNW = Q(1, -1), qSE = Q(-1, 1), qSW = Q(-1, -1);
      const T jpp = (pE - pW) * (qN - qS) - (pN - pS) * (qE - qW);
      const T jpx = pE * (qNE - qSE) - pW * (qNW - qSW) - pN * (qNE - qNW) + pS * (qSE - qSW);
      const T jxp = qN * (pNE - pNW) - qS * (pSE - pSW) - qE * (pNE - pSE) + qW * (pNW - pSW);
      out[j * Nx + i] = -((jpp + jpx + jxp) / den);
    }
}

// -div(u q) on the interior, ghost ring 0 (stencils.weno_advection).
// scratch: (ny+1)(nx+1) corner values, ny(nx+1) x-fluxes, (ny+1)nx y-fluxes.
template <class T>
void weno(const T* psi, const T* q, T* out, int Ny, int Nx, T dx, T dy, T w, T* scratch) {
  const int ny = Ny - 2, nx = Nx - 2;
  const long cw = nx + 1;
  T* C = scratch;                  // (ny+1) x (nx+1)
  T* Fx = C + (ny + 1) * cw;       // ny x (nx+1)
  T* Fy = Fx + (long)ny * cw;      // (ny+1) x nx
#pragma omp parallel for schedule(static)
  for (long b = 0; b <= ny; ++b)
    for (long a = 0; a <= nx; ++a) {
      T v = (T)0.25 * (psi[b * Nx + a] + psi[b * Nx + a + 1] + psi[(b + 1) * Nx + a] + psi[(b + 1) * Nx + a + 1]);
      if (b == 0 || b == ny || a == 0 || a == nx) v = w;
      C[b * cw + a] = v;
    }
#pragma omp parallel for schedule(static)
  for (long r = 0; r < ny; ++r) {
    const T* line = q + (r + 1) * Nx;  // line[k] = interior cell k of row r
    Fx[r * cw] = 0;
    Fx[r * cw + nx] = 0;
    for (int f = 1; f < nx; ++f) {
      const T u = -(C[(r + 1) * cw + f] - C[r * cw + f]) / dy;
      Fx[r * cw + f] = u * face_value(line, 1, nx, f, u);
    }
  }
#pragma omp parallel for schedule(static)
  for (long g = 0; g <= ny; ++g)
    for (long a = 0; a < nx; ++a) {
      T F = 0;
      if (g >= 1 && g < ny) {
        const T v = (C[g * cw + a + 1] - C[g * cw + a]) / dx;
        const T* col = q + (a + 1);  // col[k * Nx] = interior cell k of column a
        F = v * face_value(col, (long)Nx, ny, (int)g, v);
      }
      Fy[g * nx + a] = F;
    }
#pragma omp parallel for schedule(static)
  for (long j = 0; j < Ny; ++j)
    for (long i = 0; i < Nx; ++i) {
      T v = 0;
      if (j >= 1 && j <= ny && i >= 1 && i <= nx) {
        const long r = j - 1, a = i - 1;
        v = -(Fx[r * cw + a + 1] - Fx[r * cw + a]) / dx - (Fy[(r + 1) * nx + a] - Fy[r * nx + a]) / dy;
      }
      out[j * Nx + i] = v;
    }
}

template <class T>
int stencil_rhs(int nl, int Ny, int Nx, double dx_, double dy_, int advection, const T* psi, const T* q, const T* qt,
                const T* wind, const double* wind_coef, double kappa_, double nu_, double nu4_, const double* psi_wall,
                T* dq) {
  const long plane = (long)Ny * Nx;
  const T dx = (T)dx_, dy = (T)dy_, kappa = (T)kappa_, nu = (T)nu_, nu4 = (T)nu4_;
  const T dx2 = dx * dx, dy2 = dy * dy;
  std::vector<T> buf((size_t)(4 * plane + 4 * (Ny + Nx + 2)));
  T* lap = buf.data();
  T* rest = lap + plane;
  for (int l = 0; l < nl; ++l) {
    T* out = dq + l * plane;
    if (advection == 1)
      weno(psi + l * plane, qt + l * plane, out, Ny, Nx, dx, dy, (T)psi_wall[l], rest);
    else
      arakawa(psi + l * plane, qt + l * plane, out, Ny, Nx, dx, dy);
    const T wc = (T)wind_coef[l];
#pragma omp parallel for schedule(static)
    for (long k = 0; k < plane; ++k) out[k] = out[k] + wc * wind[k];
    if (kappa_ != 0.0 && l == nl - 1) {
      laplacian(psi + l * plane, lap, Ny, Nx, dx2, dy2);
#pragma omp parallel for schedule(static)
      for (long k = 0; k < plane; ++k) out[k] = out[k] + -kappa * lap[k];
    }
    if (nu_ != 0.0) {
      laplacian(q + l * plane, lap, Ny, Nx, dx2, dy2);
#pragma omp parallel for schedule(static)
      for (long k = 0; k < plane; ++k) out[k] = out[k] + nu * lap[k];
    }
    if (nu4_ != 0.0) {
      laplacian(q + l * plane, rest, Ny, Nx, dx2, dy2);
      laplacian(rest, lap, Ny, Nx, dx2, dy2);
#pragma omp parallel for schedule(static)
      for (long k = 0; k < plane; ++k) out[k] = out[k] - nu4 * lap[k];
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int qgrhs_stencil_f64(int nl, int Ny, int Nx, double dx, double dy, int advection, const double* psi, const double* q,
                      const double* qt, const double* wind, const double* wind_coef, double kappa, double nu,
                      double nu4, const double* psi_wall, double* dq) {
  return stencil_rhs<double>(nl, Ny, Nx, dx, dy, advection, psi, q, qt, wind, wind_coef, kappa, nu, nu4, psi_wall,
                             dq);
}

int qgrhs_stencil_f32(int nl, int Ny, int Nx, double dx, double dy, int advection, const float* psi, const float* q,
                      const float* qt, const float* wind, const double* wind_coef, double kappa, double nu, double nu4,
                      const double* psi_wall, float* dq) {
  return stencil_rhs<float>(nl, Ny, Nx, dx, dy, advection, psi, q, qt, wind, wind_coef, kappa, nu, nu4, psi_wall, dq);
}

// ---------------------------------------------------------------------------
// FP32 elementwise passes of the CPU baseline's step (oracle/cpu.py:FftCpuStep);
// the sine transforms between them are pocketfft's.
// ---------------------------------------------------------------------------

// qbc = q with its ghost ring set to 0; qt = qbc + beta (beta: Ny values per
// row); qm = cl2m . qbc (nl x nl, row-major).
void fftcpu_prepare(int nl, int Ny, int Nx, const float* q, const float* beta_row, const float* cl2m, float* qbc,
                    float* qt, float* qm) {
  const long plane = (long)Ny * Nx;
#pragma omp parallel for schedule(static)
  for (long j = 0; j < Ny; ++j) {
    for (long i = 0; i < Nx; ++i) {
      const bool ghost = j == 0 || j == Ny - 1 || i == 0 || i == Nx - 1;
      float v[8];
      for (int l = 0; l < nl; ++l) {
        const long k = l * plane + j * Nx + i;
        v[l] = ghost ? 0.0f : q[k];
        qbc[k] = v[l];
        qt[k] = v[l] + beta_row[j];
      }
      for (int m = 0; m < nl; ++m) {
        float a = 0.0f;
        for (int l = 0; l < nl; ++l) a += cl2m[m * nl + l] * v[l];
        qm[m * plane + j * Nx + i] = a;
      }
    }
  }
}

// f /= den elementwise over n values.
void fftcpu_divide(long n, float* f, const float* den) {
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n; ++k) f[k] = f[k] / den[k];
}

// Mass constraint and layer mix: alpha[m] = -mean(pm[m]) / hmean[m] (FP64
// means; skipped when homsol is null), psi = cm2l . (pm + alpha homsol),
// ghost ring of psi 0 when zero_ghost; wall[l] = (cm2l . alpha)[l].
void fftcpu_finish(int nl, int Ny, int Nx, const float* pm, const float* homsol, const double* hmean,
                   const float* cm2l, const double* cm2l64, int zero_ghost, float* psi, double* wall) {
  const long plane = (long)Ny * Nx;
  double alpha[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (homsol) {
    for (int m = 0; m < nl; ++m) {
      double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
      for (long k = 0; k < plane; ++k) sum += (double)pm[m * plane + k];
      alpha[m] = -(sum / (double)plane) / hmean[m];
    }
  }
  for (int l = 0; l < nl; ++l) {
    double w = 0.0;
    for (int m = 0; m < nl; ++m) w += cm2l64[l * nl + m] * alpha[m];
    wall[l] = w;
  }
  float af[8];
  for (int m = 0; m < nl; ++m) af[m] = (float)alpha[m];
#pragma omp parallel for schedule(static)
  for (long j = 0; j < Ny; ++j) {
    for (long i = 0; i < Nx; ++i) {
      const long k = j * Nx + i;
      float v[8];
      for (int m = 0; m < nl; ++m) v[m] = pm[m * plane + k] + (homsol ? af[m] * homsol[m * plane + k] : 0.0f);
      const bool ghost = j == 0 || j == Ny - 1 || i == 0 || i == Nx - 1;
      for (int l = 0; l < nl; ++l) {
        float a = 0.0f;
        for (int m = 0; m < nl; ++m) a += cm2l[l * nl + m] * v[m];
        psi[l * plane + k] = (zero_ghost && ghost) ? 0.0f : a;
      }
    }
  }
}

// out = a * x + b * (y + h * k) elementwise (the SSP-RK3 stage forms).
void fftcpu_combine(long n, float* out, float a, const float* x, float b, const float* y, float h, const float* k) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) out[i] = a * x[i] + b * (y[i] + h * k[i]);
}

void qgrhs_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

int qgrhs_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
