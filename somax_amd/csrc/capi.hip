// C ABI of the somax HIP library: plans, workspace layout and the QG entry
// points declared in include/somaxhip.h.

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/somaxhip.h"
#include "common.h"
#include "kernels.h"

namespace somax {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

static long round_up(long v, long m) { return (v + m - 1) / m * m; }

// Sine matrices of one axis (N points incl. ghosts), zero padded so MFMA
// tiles never read out of bounds: rows to a multiple of 128, pitch to a
// multiple of 32 elements.  "full" is S[n][k]; "int" is S[n][k+1] (the
// columns that meet the interior points when the ghost ring is zero).
struct SineSet {
  int N = 0;
  long ld = 0, rows = 0;
  float* f32_full = nullptr;
  float* f32_int = nullptr;
  uint16_t* hi_full = nullptr;
  uint16_t* lo_full = nullptr;
  uint16_t* hi_int = nullptr;
  uint16_t* lo_int = nullptr;
  float* lam = nullptr;  // DST-I eigenvalues (2cos(pi(k+1)/(N+1)) - 2)/d^2
};

struct DeviceBlock {
  std::vector<void*> ptrs;
  ~DeviceBlock() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** out, size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
    if (e != hipSuccess) return e;
    ptrs.push_back(p);
    *out = static_cast<T*>(p);
    return hipSuccess;
  }
};

static hipError_t build_sines(DeviceBlock& mem, SineSet& s, int N, double d) {
  s.N = N;
  s.ld = round_up(N, 32);
  s.rows = round_up(N, 128);
  const size_t count = (size_t)s.rows * s.ld;
  std::vector<float> ff(count, 0.0f), fi(count, 0.0f);
  std::vector<uint16_t> hf(count, 0), lf(count, 0), hi(count, 0), li(count, 0);
  const double c = std::sqrt(2.0 / (N + 1));
  for (int n = 0; n < N; ++n)
    for (int k = 0; k < N; ++k) {
      const double v = c * std::sin(M_PI * (double)(n + 1) * (double)(k + 1) / (double)(N + 1));
      const float f = (float)v;
      const uint16_t h = f32_to_bf16_bits(f);
      const uint16_t l = f32_to_bf16_bits((float)(v - (double)bf16_bits_to_f32(h)));
      ff[(size_t)n * s.ld + k] = f;
      hf[(size_t)n * s.ld + k] = h;
      lf[(size_t)n * s.ld + k] = l;
      if (k >= 1 && k <= N - 2) {
        fi[(size_t)n * s.ld + k - 1] = f;
        hi[(size_t)n * s.ld + k - 1] = h;
        li[(size_t)n * s.ld + k - 1] = l;
      }
    }
  std::vector<float> lam(N);
  for (int k = 0; k < N; ++k) lam[k] = (float)((2.0 * std::cos(M_PI * (k + 1) / (double)(N + 1)) - 2.0) / (d * d));
  hipError_t e;
#define SOMAX_UP(dst, src, n)                                                            \
  if ((e = mem.alloc(&dst, n)) != hipSuccess) return e;                                 \
  if ((e = hipMemcpy(dst, src.data(), (n) * sizeof(src[0]), hipMemcpyHostToDevice)) != \
      hipSuccess)                                                                        \
    return e;
  SOMAX_UP(s.f32_full, ff, count);
  SOMAX_UP(s.f32_int, fi, count);
  SOMAX_UP(s.hi_full, hf, count);
  SOMAX_UP(s.lo_full, lf, count);
  SOMAX_UP(s.hi_int, hi, count);
  SOMAX_UP(s.lo_int, li, count);
  SOMAX_UP(s.lam, lam, (size_t)N);
#undef SOMAX_UP
  return hipSuccess;
}

}  // namespace somax

using namespace somax;

struct somaxhip_qg_plan_s {
  int nx, ny, nl, Nx, Ny;
  double dx, dy;
  int zero_psi_ghost, advection, precision, mass_correction;
  std::vector<double> cl2m, cm2l, lambdas, wind_scale;
  DeviceBlock mem;
  SineSet sx, sy;
  float* beta_row = nullptr;
  float* wind_row = nullptr;
  float* homsol = nullptr;
  std::vector<double> homsol_mean;
  // workspace layout (bytes offsets), all float planes
  long wplane = 0;  // elements per spectral/work plane
  long wld = 0;     // pitch of work planes
  long splane = 0;  // Ny*Nx
  size_t off_T = 0, off_Z = 0, off_V = 0, off_P0 = 0, off_R = 0, off_Q1 = 0, off_Q2 = 0, off_PSI = 0;
  size_t ws_bytes = 0;
};

extern "C" {

size_t somaxhip_last_error(char* buf, size_t len) {
  const std::string& m = g_last_error;
  if (buf && len) {
    const size_t n = m.size() < len - 1 ? m.size() : len - 1;
    std::memcpy(buf, m.data(), n);
    buf[n] = '\0';
  }
  return m.size();
}

int somaxhip_version(void) { return 100; }

int somaxhip_init(void) {
  int dev = 0;
  SOMAX_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SOMAX_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("somaxhip is built for gfx950; current device is ") + prop.gcnArchName);
    return SOMAXHIP_ERR_DEVICE;
  }
  return SOMAXHIP_OK;
}

int somaxhip_qg_plan_create(const somaxhip_qg_desc* d, somaxhip_qg_plan* out) {
  if (!d || !out) {
    set_error("somaxhip_qg_plan_create: null argument");
    return SOMAXHIP_ERR_ARG;
  }
  *out = nullptr;
  if (d->nx < 2 || d->ny < 2 || d->nl < 1 || d->nl > MAX_LAYERS || !(d->dx > 0) || !(d->dy > 0)) {
    set_error("somaxhip_qg_plan_create: need nx, ny >= 2, 1 <= nl <= 8, dx, dy > 0");
    return SOMAXHIP_ERR_ARG;
  }
  if (!d->cl2m || !d->cm2l || !d->lambdas || !d->beta_y_row || !d->wind_row || !d->wind_scale) {
    set_error("somaxhip_qg_plan_create: a required table pointer is null");
    return SOMAXHIP_ERR_ARG;
  }
  if (d->advection != SOMAXHIP_ADV_ARAKAWA && d->advection != SOMAXHIP_ADV_WENO) {
    set_error("somaxhip_qg_plan_create: unknown advection");
    return SOMAXHIP_ERR_ARG;
  }
  if (d->precision != SOMAXHIP_PREC_FP32 && d->precision != SOMAXHIP_PREC_MIXED) {
    set_error("somaxhip_qg_plan_create: unknown precision");
    return SOMAXHIP_ERR_ARG;
  }
  if (d->advection == SOMAXHIP_ADV_WENO) {
    set_error("somaxhip_qg_plan_create: WENO advection is not available in this build");
    return SOMAXHIP_ERR_ARG;
  }
  if (d->mass_correction) {
    set_error("somaxhip_qg_plan_create: mass correction is not available in this build");
    return SOMAXHIP_ERR_ARG;
  }
  if (d->mass_correction && (!d->homsol || !d->homsol_mean)) {
    set_error("somaxhip_qg_plan_create: mass_correction needs homsol and homsol_mean");
    return SOMAXHIP_ERR_ARG;
  }
  somaxhip_qg_plan p = new (std::nothrow) somaxhip_qg_plan_s();
  if (!p) {
    set_error("somaxhip_qg_plan_create: out of host memory");
    return SOMAXHIP_ERR_ALLOC;
  }
  p->nx = d->nx;
  p->ny = d->ny;
  p->nl = d->nl;
  p->Nx = d->nx + 2;
  p->Ny = d->ny + 2;
  p->dx = d->dx;
  p->dy = d->dy;
  p->zero_psi_ghost = d->zero_psi_ghost ? 1 : 0;
  p->advection = d->advection;
  p->precision = d->precision;
  p->mass_correction = d->mass_correction ? 1 : 0;
  const int nl = d->nl;
  p->cl2m.assign(d->cl2m, d->cl2m + nl * nl);
  p->cm2l.assign(d->cm2l, d->cm2l + nl * nl);
  p->lambdas.assign(d->lambdas, d->lambdas + nl);
  p->wind_scale.assign(d->wind_scale, d->wind_scale + nl);

  hipError_t e;
  if ((e = build_sines(p->mem, p->sx, p->Nx, p->dx)) != hipSuccess ||
      (e = build_sines(p->mem, p->sy, p->Ny, p->dy)) != hipSuccess) {
    set_error(std::string("somaxhip_qg_plan_create: ") + hipGetErrorString(e));
    delete p;
    return SOMAXHIP_ERR_ALLOC;
  }
  std::vector<float> beta(p->Ny), wind(p->Ny);
  for (int j = 0; j < p->Ny; ++j) {
    beta[j] = (float)d->beta_y_row[j];
    wind[j] = (float)d->wind_row[j];
  }
  if ((e = p->mem.alloc(&p->beta_row, p->Ny)) != hipSuccess ||
      (e = hipMemcpy(p->beta_row, beta.data(), p->Ny * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = p->mem.alloc(&p->wind_row, p->Ny)) != hipSuccess ||
      (e = hipMemcpy(p->wind_row, wind.data(), p->Ny * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
    set_error(std::string("somaxhip_qg_plan_create: ") + hipGetErrorString(e));
    delete p;
    return SOMAXHIP_ERR_ALLOC;
  }
  if (p->mass_correction) {
    const size_t cnt = (size_t)nl * p->Ny * p->Nx;
    if ((e = p->mem.alloc(&p->homsol, cnt)) != hipSuccess ||
        (e = hipMemcpy(p->homsol, d->homsol, cnt * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
      set_error(std::string("somaxhip_qg_plan_create: ") + hipGetErrorString(e));
      delete p;
      return SOMAXHIP_ERR_ALLOC;
    }
    p->homsol_mean.assign(d->homsol_mean, d->homsol_mean + nl);
  }

  // workspace layout
  const long maxN = p->Nx > p->Ny ? p->Nx : p->Ny;
  p->wld = round_up(maxN, 64);
  p->wplane = maxN * p->wld;
  p->splane = (long)p->Ny * p->Nx;
  size_t off = 0;
  auto take = [&](size_t floats) {
    const size_t o = off;
    off += round_up((long)(floats * sizeof(float)), 256);
    return o;
  };
  p->off_T = take((size_t)nl * p->wplane);
  p->off_Z = take((size_t)nl * p->wplane);
  p->off_V = take((size_t)nl * p->wplane);
  if (p->precision == SOMAXHIP_PREC_MIXED) {
    p->off_P0 = take((size_t)nl * p->splane);
    p->off_R = take((size_t)nl * p->splane);
  }
  p->off_Q1 = take((size_t)nl * p->splane);
  p->off_Q2 = take((size_t)nl * p->splane);
  p->off_PSI = take((size_t)nl * p->splane);
  p->ws_bytes = off;
  *out = p;
  return SOMAXHIP_OK;
}

int somaxhip_qg_plan_destroy(somaxhip_qg_plan p) {
  delete p;
  return SOMAXHIP_OK;
}

size_t somaxhip_qg_workspace_bytes(somaxhip_qg_plan p) { return p ? p->ws_bytes : 0; }

}  // extern "C"

namespace {

inline float* wsf(void* ws, size_t off) { return reinterpret_cast<float*>(static_cast<char*>(ws) + off); }

// Optional per-launch timing: when a profile is active on this thread every
// launch below is bracketed by a pair of HIP events on its stream.
struct ProfEntry {
  const char* name;
  double flops, bytes;
  hipEvent_t a, b;
};
struct Profile {
  std::vector<ProfEntry> entries;
};
thread_local Profile* g_prof = nullptr;

struct ProfScope {
  hipStream_t st;
  bool on;
  ProfScope(const char* name, double flops, double bytes, hipStream_t s) : st(s), on(g_prof != nullptr) {
    if (!on) return;
    ProfEntry e{name, flops, bytes, nullptr, nullptr};
    (void)hipEventCreate(&e.a);
    (void)hipEventCreate(&e.b);
    (void)hipEventRecord(e.a, st);
    g_prof->entries.push_back(e);
  }
  ~ProfScope() {
    if (on) (void)hipEventRecord(g_prof->entries.back().b, st);
  }
};

const char* gemm_name(int prec) {
  return prec == GP_FP32 ? "dst_gemm_fp32" : (prec == GP_BF16X3 ? "dst_gemm_bf16x3" : "dst_gemm_bf16x1");
}

void gemm(const GemmArgs& a, int prec, bool trans, int epi, int batch, hipStream_t st) {
  // algorithmic work: 2*R*C*K flops; bytes: X read + O written once
  const double flops = 2.0 * a.R * (double)a.C * a.K * batch;
  const double bytes = 4.0 * batch * ((double)a.R * a.K + (double)a.R * a.C);
  ProfScope ps(gemm_name(prec), flops, bytes, st);
  launch_dst_gemm(a, prec, trans, epi, batch, st);
}

// One DST-I layer-space solve of (lap_h - B) psi = x.
//   interior_input: the ghost ring of x is zero (q with its boundary
//   condition), so the forward transform contracts over interior points only.
//   out: psi in state layout (ld Nx); Oadd/mask as in GE_FINAL.
void dst_solve(somaxhip_qg_plan p, const float* x, long x_bs, bool interior_input, int prec_fwd,
               int prec_inv, float* out, const float* addend, bool mask, void* ws, hipStream_t st) {
  float* T = wsf(ws, p->off_T);
  float* Z = wsf(ws, p->off_Z);
  float* V = wsf(ws, p->off_V);
  const int Nx = p->Nx, Ny = p->Ny, nl = p->nl;

  // F1: T[xm][y] = sum_x X[y][x] Sx[xm][x]
  GemmArgs g{};
  g.bso = p->wplane;
  g.ldo = p->wld;
  if (interior_input) {
    g.X = x + Nx + 1;
    g.R = Ny - 2;
    g.K = Nx - 2;
    g.Sf = p->sx.f32_int;
    g.Sh = p->sx.hi_int;
    g.Sl = p->sx.lo_int;
  } else {
    g.X = x;
    g.R = Ny;
    g.K = Nx;
    g.Sf = p->sx.f32_full;
    g.Sh = p->sx.hi_full;
    g.Sl = p->sx.lo_full;
  }
  g.ldx = Nx;
  g.bsx = x_bs;
  g.lds = p->sx.ld;
  g.C = Nx;
  g.O = T;
  gemm(g, prec_fwd, true, GE_STORE, nl, st);

  // F2: Z[xm][ym] = sum_y T[xm][y] Sy[ym][y]
  GemmArgs h{};
  h.X = T;
  h.ldx = p->wld;
  h.bsx = p->wplane;
  h.R = Nx;
  h.K = interior_input ? Ny - 2 : Ny;
  h.Sf = interior_input ? p->sy.f32_int : p->sy.f32_full;
  h.Sh = interior_input ? p->sy.hi_int : p->sy.hi_full;
  h.Sl = interior_input ? p->sy.lo_int : p->sy.lo_full;
  h.lds = p->sy.ld;
  h.C = Ny;
  h.O = Z;
  h.ldo = p->wld;
  h.bso = p->wplane;
  gemm(h, prec_fwd, false, GE_STORE, nl, st);

  // spectral operator
  SpectralArgs s{};
  s.Z = Z;
  s.ld = p->wld;
  s.bs = p->wplane;
  s.nxm = Nx;
  s.nym = Ny;
  s.nl = nl;
  s.lamx = p->sx.lam;
  s.lamy = p->sy.lam;
  for (int a = 0; a < nl; ++a) {
    s.lam[a] = (float)p->lambdas[a];
    for (int b = 0; b < nl; ++b) {
      s.cl2m[a * MAX_LAYERS + b] = (float)p->cl2m[a * nl + b];
      s.cm2l[a * MAX_LAYERS + b] = (float)p->cm2l[a * nl + b];
    }
  }
  {
    ProfScope ps("spectral", 0.0, 8.0 * nl * (double)Nx * Ny, st);
    launch_spectral(s, st);
  }

  // I1: V[y][xm] = sum_ym Sy[y][ym] W[xm][ym]
  GemmArgs u{};
  u.X = Z;
  u.ldx = p->wld;
  u.bsx = p->wplane;
  u.R = Nx;
  u.K = Ny;
  u.Sf = p->sy.f32_full;
  u.Sh = p->sy.hi_full;
  u.Sl = p->sy.lo_full;
  u.lds = p->sy.ld;
  u.C = Ny;
  u.O = V;
  u.ldo = p->wld;
  u.bso = p->wplane;
  gemm(u, prec_inv, true, GE_STORE, nl, st);

  // I2: psi[y][x] = sum_xm V[y][xm] Sx[x][xm]
  GemmArgs w{};
  w.X = V;
  w.ldx = p->wld;
  w.bsx = p->wplane;
  w.R = Ny;
  w.K = Nx;
  w.Sf = p->sx.f32_full;
  w.Sh = p->sx.hi_full;
  w.Sl = p->sx.lo_full;
  w.lds = p->sx.ld;
  w.C = Nx;
  w.O = out;
  w.ldo = Nx;
  w.bso = p->splane;
  w.Oadd = addend;
  w.mask_ghost = mask ? 1 : 0;
  w.phys_rows = Ny;
  w.phys_cols = Nx;
  gemm(w, prec_inv, false, GE_FINAL, nl, st);
}

void invert(somaxhip_qg_plan p, const float* q, float* psi, void* ws, hipStream_t st, bool apply_bc = true) {
  const bool mask = p->zero_psi_ghost != 0;
  if (p->precision == SOMAXHIP_PREC_FP32) {
    dst_solve(p, q, p->splane, apply_bc, GP_FP32, GP_FP32, psi, nullptr, mask, ws, st);
    return;
  }
  // mixed: bf16x3 solve, FP32 Helmholtz residual, bf16 correction solve
  float* P0 = wsf(ws, p->off_P0);
  float* R = wsf(ws, p->off_R);
  dst_solve(p, q, p->splane, apply_bc, GP_BF16X3, GP_BF16X3, P0, nullptr, false, ws, st);
  ResidualArgs r{};
  r.q = q;
  r.q_ld = p->Nx;
  r.q_bs = p->splane;
  r.psi = P0;
  r.p_ld = p->Nx;
  r.p_bs = p->splane;
  r.r = R;
  r.r_ld = p->Nx;
  r.r_bs = p->splane;
  r.Nx = p->Nx;
  r.Ny = p->Ny;
  r.nl = p->nl;
  r.apply_bc = apply_bc ? 1 : 0;
  r.idx2 = (float)(1.0 / (p->dx * p->dx));
  r.idy2 = (float)(1.0 / (p->dy * p->dy));
  const int nl = p->nl;
  for (int l = 0; l < nl; ++l)
    for (int k = 0; k < nl; ++k) {
      double b = 0.0;
      for (int m = 0; m < nl; ++m) b += p->cm2l[l * nl + m] * p->lambdas[m] * p->cl2m[m * nl + k];
      r.B[l * MAX_LAYERS + k] = (float)b;
    }
  {
    ProfScope ps("residual", 0.0, 12.0 * nl * (double)p->splane, st);
    launch_residual(r, st);
  }
  dst_solve(p, R, p->splane, false, GP_BF16X3, GP_BF16X1, psi, P0, mask, ws, st);
}

void stage(somaxhip_qg_plan p, const float* qs, const float* psi, const float* qn, float* out,
           const somaxhip_qg_params* prm, float c0, float c1, float c2, bool apply_bc, hipStream_t st) {
  StageArgs a{};
  a.qs = qs;
  a.psi = psi;
  a.qn = qn;
  a.out = out;
  a.Nx = p->Nx;
  a.Ny = p->Ny;
  a.nl = p->nl;
  a.plane = p->splane;
  a.beta_row = p->beta_row;
  a.wind_row = p->wind_row;
  for (int l = 0; l < p->nl; ++l) a.wind_coef[l] = (float)(prm->tau0 * p->wind_scale[l]);
  a.nu = (float)prm->nu;
  a.nu4 = (float)prm->nu4;
  a.kappa = (float)prm->kappa;
  a.idx2 = (float)(1.0 / (p->dx * p->dx));
  a.idy2 = (float)(1.0 / (p->dy * p->dy));
  a.iarakawa = (float)(1.0 / (12.0 * p->dx * p->dy));
  a.idx = (float)(1.0 / p->dx);
  a.idy = (float)(1.0 / p->dy);
  a.c0 = c0;
  a.c1 = c1;
  a.c2 = c2;
  a.advection = p->advection;
  a.apply_bc = apply_bc ? 1 : 0;
  // algorithmic bytes: read qs and psi, read qn when combined, write out
  const double bytes = 4.0 * p->nl * (double)p->nx * p->ny * (3.0 + (c0 != 0.0f ? 1.0 : 0.0));
  ProfScope ps("qg_stage", 0.0, bytes, st);
  launch_qg_stage(a, st);
}

int check_call(somaxhip_qg_plan p, const void* a, const void* b, void* ws, const char* who) {
  if (!p || !a || !b || !ws) {
    set_error(std::string(who) + ": null plan, buffer or workspace");
    return SOMAXHIP_ERR_ARG;
  }
  return SOMAXHIP_OK;
}

}  // namespace

extern "C" {

int somaxhip_qg_invert(somaxhip_qg_plan p, const float* q, float* psi, int apply_bc, void* ws,
                       somaxhip_stream_t stream) {
  int rc = check_call(p, q, psi, ws, "somaxhip_qg_invert");
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  invert(p, q, psi, ws, st, apply_bc != 0);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_rhs(somaxhip_qg_plan p, const float* q, float* dq, float* psi_out,
                    const somaxhip_qg_params* prm, int apply_bc, void* ws, somaxhip_stream_t stream) {
  int rc = check_call(p, q, dq, ws, "somaxhip_qg_rhs");
  if (rc) return rc;
  if (!prm) {
    set_error("somaxhip_qg_rhs: null params");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float* psi = psi_out ? psi_out : wsf(ws, p->off_PSI);
  invert(p, q, psi, ws, st, apply_bc != 0);
  stage(p, q, psi, nullptr, dq, prm, 0.0f, 0.0f, 1.0f, apply_bc != 0, st);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_step(somaxhip_qg_plan p, float* q, float* psi_out, int nsteps, double dt,
                     const somaxhip_qg_params* prm, void* ws, somaxhip_stream_t stream) {
  int rc = check_call(p, q, q, ws, "somaxhip_qg_step");
  if (rc) return rc;
  if (!prm || nsteps < 0) {
    set_error("somaxhip_qg_step: null params or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float* psi = wsf(ws, p->off_PSI);
  float* q1 = wsf(ws, p->off_Q1);
  float* q2 = wsf(ws, p->off_Q2);
  const float h = (float)dt;
  for (int n = 0; n < nsteps; ++n) {
    invert(p, q, psi, ws, st);
    stage(p, q, psi, nullptr, q1, prm, 0.0f, 1.0f, h, true, st);
    invert(p, q1, psi, ws, st);
    stage(p, q1, psi, q, q2, prm, 0.75f, 0.25f, 0.25f * h, true, st);
    invert(p, q2, psi, ws, st);
    stage(p, q2, psi, q, q, prm, 1.0f / 3.0f, 2.0f / 3.0f, 2.0f / 3.0f * h, true, st);
  }
  if (psi_out && nsteps > 0)
    SOMAX_HIP_CHECK(hipMemcpyAsync(psi_out, psi, (size_t)p->nl * p->splane * sizeof(float),
                                   hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

}  // extern "C"

extern "C" int somaxhip_zero_ghost(float* f, int nl, int ny, int nx, somaxhip_stream_t stream) {
  if (!f || nl < 1 || nx < 1 || ny < 1) {
    set_error("somaxhip_zero_ghost: null field or bad shape");
    return SOMAXHIP_ERR_ARG;
  }
  launch_zero_ghost(f, nl, ny + 2, nx + 2, reinterpret_cast<hipStream_t>(stream));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

extern "C" int somaxhip_qg_step_profile(somaxhip_qg_plan p, float* q, int nsteps, double dt,
                                        const somaxhip_qg_params* prm, void* ws, somaxhip_stream_t stream,
                                        somaxhip_kernel_stat* stats, int max_stats, int* nstats) {
  if (!stats || !nstats || max_stats < 1) {
    set_error("somaxhip_qg_step_profile: null stats");
    return SOMAXHIP_ERR_ARG;
  }
  Profile prof;
  g_prof = &prof;
  const int rc = somaxhip_qg_step(p, q, nullptr, nsteps, dt, prm, ws, stream);
  g_prof = nullptr;
  hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
  int n = 0;
  for (auto& en : prof.entries) {
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, en.a, en.b);
    (void)hipEventDestroy(en.a);
    (void)hipEventDestroy(en.b);
    int k = 0;
    while (k < n && std::strcmp(stats[k].name, en.name) != 0) ++k;
    if (k == n) {
      if (n == max_stats) continue;
      std::memset(&stats[n], 0, sizeof(stats[n]));
      std::strncpy(stats[n].name, en.name, sizeof(stats[n].name) - 1);
      ++n;
    }
    stats[k].launches += 1;
    stats[k].total_ms += ms;
    stats[k].flops += en.flops;
    stats[k].bytes += en.bytes;
  }
  *nstats = n;
  if (rc) return rc;
  SOMAX_HIP_CHECK(e);
  return SOMAXHIP_OK;
}
