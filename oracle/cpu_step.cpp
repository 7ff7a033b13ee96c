// CPU restatement of the multilayer QG SSP-RK3 step in C++ with OpenMP.
//
// TEST INFRASTRUCTURE ONLY: a CPU baseline timed beside the GPU (bench.py's
// cpu_baseline leg) and checked against the NumPy oracle (oracle/qg.py, which
// restates somax/_src/models/qg/baroclinic.py:135-195); the product never
// loads it.  Same algorithm as oracle/qg.py's Arakawa path without the WENO
// and mass-correction extensions:
//   psi = Cm2l . DST-I^-1[ DST-I(Cl2m . q) / (Dy + Dx - lambda) ]   (dense
//         orthonormal sine matrices over the full (ny+2) x (nx+2) array;
//         ghost ring of psi set to 0 when zero_psi_ghost)
//   dq  = -J(psi, q + beta_y) on the interior (Arakawa 1966)
//         + tau0 * wind_scale[l] * wind_row[j] over the whole array
//         - kappa * lap(psi[nl-1]) + nu * lap(q) - nu4 * lap(lap(q))
// with q read with its ghost ring set to 0 (the boundary condition before
// every right-hand side) and the SSP-RK3 update in Shu-Osher form.  The real
// type is float or double.  Loops are parallel over rows (OpenMP).

#include <omp.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

namespace {

template <class T>
struct Plan {
  int nx, ny, nl, Nx, Ny;
  double dx, dy;
  std::vector<T> Sx, Sy;      // orthonormal DST-I matrices (Nx x Nx, Ny x Ny), symmetric
  std::vector<T> den;         // nl x Ny x Nx: Dy[j] + Dx[i] - lambda_m
  std::vector<double> cl2m, cm2l;
  std::vector<T> beta_y, wind;  // Ny rows each
  std::vector<T> wscale;
  int zero_psi_ghost;
  T nu, nu4, kappa, tau0;
  // scratch
  std::vector<T> qb, qm, tmp, psi, lap, dq, q1, q2;
};

template <class T>
void sine(std::vector<T>& S, int n) {
  S.resize((size_t)n * n);
  const double c = std::sqrt(2.0 / (n + 1));
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) S[(size_t)j * n + k] = (T)(c * std::sin(M_PI * (j + 1) * (k + 1) / (n + 1)));
}

// C[M][N] = A[M][K] . B[K][N], row-major, parallel over rows of C.
template <class T>
void matmul(const T* A, const T* B, T* C, int M, int K, int N) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; ++i) {
    T* c = C + (size_t)i * N;
    for (int j = 0; j < N; ++j) c[j] = 0;
    for (int k = 0; k < K; ++k) {
      const T a = A[(size_t)i * K + k];
      const T* b = B + (size_t)k * N;
#pragma omp simd
      for (int j = 0; j < N; ++j) c[j] += a * b[j];
    }
  }
}

template <class T>
void invert(Plan<T>& p, const T* q, T* psi) {
  const int Ny = p.Ny, Nx = p.Nx, nl = p.nl;
  const size_t pl = (size_t)Ny * Nx;
  // layer -> mode
#pragma omp parallel for schedule(static)
  for (long s = 0; s < (long)pl; ++s)
    for (int m = 0; m < nl; ++m) {
      double acc = 0.0;
      for (int l = 0; l < nl; ++l) acc += p.cl2m[m * nl + l] * (double)q[l * pl + s];
      p.qm[m * pl + s] = (T)acc;
    }
  for (int m = 0; m < nl; ++m) {
    T* x = p.qm.data() + m * pl;
    matmul(p.Sy.data(), x, p.tmp.data(), Ny, Ny, Nx);  // Sy . X
    matmul(p.tmp.data(), p.Sx.data(), x, Ny, Nx, Nx);  // . Sx
    const T* d = p.den.data() + m * pl;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)pl; ++s) x[s] /= d[s];
    matmul(p.Sy.data(), x, p.tmp.data(), Ny, Ny, Nx);
    matmul(p.tmp.data(), p.Sx.data(), x, Ny, Nx, Nx);
  }
  // mode -> layer
#pragma omp parallel for schedule(static)
  for (long s = 0; s < (long)pl; ++s) {
    const int j = (int)(s / Nx), i = (int)(s % Nx);
    const bool ghost = p.zero_psi_ghost && (j == 0 || i == 0 || j == Ny - 1 || i == Nx - 1);
    for (int l = 0; l < nl; ++l) {
      double acc = 0.0;
      for (int m = 0; m < nl; ++m) acc += p.cm2l[l * nl + m] * (double)p.qm[m * pl + s];
      psi[l * pl + s] = ghost ? (T)0 : (T)acc;
    }
  }
}

template <class T>
inline T lap5(const T* f, long s, int Nx, T idx2, T idy2) {
  return (f[s + 1] - 2 * f[s] + f[s - 1]) * idx2 + (f[s + Nx] - 2 * f[s] + f[s - Nx]) * idy2;
}

// dq = RHS(BC(q)).
template <class T>
void rhs(Plan<T>& p, const T* q, T* dq) {
  const int Ny = p.Ny, Nx = p.Nx, nl = p.nl;
  const size_t pl = (size_t)Ny * Nx;
  T* qb = p.qb.data();
#pragma omp parallel for schedule(static)
  for (long s = 0; s < (long)(nl * pl); ++s) {
    const long r = s % (long)pl;
    const int j = (int)(r / Nx), i = (int)(r % Nx);
    qb[s] = (j == 0 || i == 0 || j == Ny - 1 || i == Nx - 1) ? (T)0 : q[s];
  }
  invert(p, qb, p.psi.data());
  const T idx2 = (T)(1.0 / (p.dx * p.dx)), idy2 = (T)(1.0 / (p.dy * p.dy));
  const T jc = (T)(1.0 / (12.0 * p.dx * p.dy));
  T* lap = p.lap.data();
  // inner Laplacian of q (ghost ring 0) for the biharmonic term
#pragma omp parallel for collapse(2) schedule(static)
  for (int l = 0; l < nl; ++l)
    for (int j = 0; j < Ny; ++j)
      for (int i = 0; i < Nx; ++i) {
        const long s = l * pl + (long)j * Nx + i;
        lap[s] = (j == 0 || i == 0 || j == Ny - 1 || i == Nx - 1) ? (T)0 : lap5(qb, s, Nx, idx2, idy2);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int l = 0; l < nl; ++l)
    for (int j = 0; j < Ny; ++j) {
      const T w = p.tau0 * p.wscale[l] * p.wind[j];
      for (int i = 0; i < Nx; ++i) {
        const long s = l * pl + (long)j * Nx + i;
        T v = w;
        if (!(j == 0 || i == 0 || j == Ny - 1 || i == Nx - 1)) {
          const T* ps = p.psi.data() + l * pl;
          const long c = (long)j * Nx + i;
          auto P = [&](int dj, int di) { return ps[c + (long)dj * Nx + di]; };
          auto Q = [&](int dj, int di) { return qb[l * pl + c + (long)dj * Nx + di] + p.beta_y[j + dj]; };
          const T jpp = (P(0, 1) - P(0, -1)) * (Q(1, 0) - Q(-1, 0)) - (P(1, 0) - P(-1, 0)) * (Q(0, 1) - Q(0, -1));
          const T jpx = P(0, 1) * (Q(1, 1) - Q(-1, 1)) - P(0, -1) * (Q(1, -1) - Q(-1, -1)) -
                        P(1, 0) * (Q(1, 1) - Q(1, -1)) + P(-1, 0) * (Q(-1, 1) - Q(-1, -1));
          const T jxp = Q(1, 0) * (P(1, 1) - P(1, -1)) - Q(-1, 0) * (P(-1, 1) - P(-1, -1)) -
                        Q(0, 1) * (P(1, 1) - P(-1, 1)) + Q(0, -1) * (P(1, -1) - P(-1, -1));
          v -= (jpp + jpx + jxp) * jc;
          if (p.kappa != 0 && l == nl - 1) v -= p.kappa * lap5(ps, c, Nx, idx2, idy2);
          if (p.nu != 0) v += p.nu * lap5(qb, s, Nx, idx2, idy2);
          if (p.nu4 != 0) v -= p.nu4 * lap5(lap, s, Nx, idx2, idy2);
        }
        dq[s] = v;
      }
    }
}

template <class T>
void step(Plan<T>& p, T* q, int nsteps, T dt) {
  const long n = (long)p.nl * p.Ny * p.Nx;
  T* d = p.dq.data();
  T* q1 = p.q1.data();
  T* q2 = p.q2.data();
  for (int k = 0; k < nsteps; ++k) {
    rhs(p, q, d);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < n; ++s) q1[s] = q[s] + dt * d[s];
    rhs(p, q1, d);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < n; ++s) q2[s] = (T)0.75 * q[s] + (T)0.25 * (q1[s] + dt * d[s]);
    rhs(p, q2, d);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < n; ++s) q[s] = q[s] / (T)3 + (T)2 / (T)3 * (q2[s] + dt * d[s]);
  }
}

template <class T>
Plan<T>* create(int nx, int ny, int nl, double dx, double dy, const double* cl2m, const double* cm2l,
                const double* lambdas, const double* beta_y, const double* wind, const double* wscale,
                int zero_psi_ghost, double nu, double nu4, double kappa, double tau0) {
  auto* p = new Plan<T>();
  p->nx = nx;
  p->ny = ny;
  p->nl = nl;
  p->Nx = nx + 2;
  p->Ny = ny + 2;
  p->dx = dx;
  p->dy = dy;
  const int Nx = p->Nx, Ny = p->Ny;
  const size_t pl = (size_t)Ny * Nx;
  sine(p->Sx, Nx);
  sine(p->Sy, Ny);
  p->den.resize(nl * pl);
  for (int m = 0; m < nl; ++m)
    for (int j = 0; j < Ny; ++j)
      for (int i = 0; i < Nx; ++i) {
        const double ey = (2.0 * std::cos(M_PI * (j + 1) / (Ny + 1)) - 2.0) / (dy * dy);
        const double ex = (2.0 * std::cos(M_PI * (i + 1) / (Nx + 1)) - 2.0) / (dx * dx);
        p->den[m * pl + (size_t)j * Nx + i] = (T)(ey + ex - lambdas[m]);
      }
  p->cl2m.assign(cl2m, cl2m + nl * nl);
  p->cm2l.assign(cm2l, cm2l + nl * nl);
  p->beta_y.resize(Ny);
  p->wind.resize(Ny);
  for (int j = 0; j < Ny; ++j) {
    p->beta_y[j] = (T)beta_y[j];
    p->wind[j] = (T)wind[j];
  }
  p->wscale.resize(nl);
  for (int l = 0; l < nl; ++l) p->wscale[l] = (T)wscale[l];
  p->zero_psi_ghost = zero_psi_ghost;
  p->nu = (T)nu;
  p->nu4 = (T)nu4;
  p->kappa = (T)kappa;
  p->tau0 = (T)tau0;
  for (auto* v : {&p->qb, &p->qm, &p->psi, &p->lap, &p->dq, &p->q1, &p->q2}) v->resize(nl * pl);
  p->tmp.resize(pl);
  return p;
}

struct Handle {
  int fp64;
  void* plan;
};

}  // namespace

extern "C" {

void* qgcpu_create(int nx, int ny, int nl, double dx, double dy, const double* cl2m, const double* cm2l,
                   const double* lambdas, const double* beta_y_rows, const double* wind_rows,
                   const double* wind_scale, int zero_psi_ghost, double nu, double nu4, double kappa, double tau0,
                   int fp64) {
  auto* h = new Handle{fp64, nullptr};
  if (fp64)
    h->plan = create<double>(nx, ny, nl, dx, dy, cl2m, cm2l, lambdas, beta_y_rows, wind_rows, wind_scale,
                             zero_psi_ghost, nu, nu4, kappa, tau0);
  else
    h->plan = create<float>(nx, ny, nl, dx, dy, cl2m, cm2l, lambdas, beta_y_rows, wind_rows, wind_scale,
                            zero_psi_ghost, nu, nu4, kappa, tau0);
  return h;
}

// q: float (fp32 plan) or double (fp64 plan), nl x (ny+2) x (nx+2), in place.
void qgcpu_step(void* handle, void* q, int nsteps, double dt) {
  auto* h = static_cast<Handle*>(handle);
  if (h->fp64)
    step(*static_cast<Plan<double>*>(h->plan), static_cast<double*>(q), nsteps, dt);
  else
    step(*static_cast<Plan<float>*>(h->plan), static_cast<float*>(q), nsteps, (float)dt);
}

void qgcpu_destroy(void* handle) {
  auto* h = static_cast<Handle*>(handle);
  if (!h) return;
  if (h->fp64)
    delete static_cast<Plan<double>*>(h->plan);
  else
    delete static_cast<Plan<float>*>(h->plan);
  delete h;
}

int qgcpu_max_threads(void) { return omp_get_max_threads(); }
void qgcpu_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
