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
// columns that meet the interior points when the ghost ring is zero).  The
// bf16 copies hold three pieces S = s0 + s1 + s2 (piece stride rows*ld).
struct SineSet {
  int N = 0;
  long ld = 0, rows = 0;
  float* f32_full = nullptr;
  float* f32_int = nullptr;
  uint16_t* b_full = nullptr;
  uint16_t* b_int = nullptr;
  float* lam = nullptr;  // DST-I eigenvalues (2cos(pi(k+1)/(N+1)) - 2)/d^2
  double lam_min = 0.0, lam_max = 0.0;  // smallest / largest |eigenvalue|
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

static void split3(double v, uint16_t* p, long stride) {
  const uint16_t a = f32_to_bf16_bits((float)v);
  const double r1 = v - (double)bf16_bits_to_f32(a);
  const uint16_t b = f32_to_bf16_bits((float)r1);
  const double r2 = r1 - (double)bf16_bits_to_f32(b);
  p[0] = a;
  p[stride] = b;
  p[2 * stride] = f32_to_bf16_bits((float)r2);
}

static hipError_t build_sines(DeviceBlock& mem, SineSet& s, int N, double d) {
  s.N = N;
  s.ld = round_up(N, 32);
  s.rows = round_up(N, 128);
  const size_t count = (size_t)s.rows * s.ld;
  std::vector<float> ff(count, 0.0f), fi(count, 0.0f);
  std::vector<uint16_t> bf(3 * count, 0), bi(3 * count, 0);
  const double c = std::sqrt(2.0 / (N + 1));
  for (int n = 0; n < N; ++n)
    for (int k = 0; k < N; ++k) {
      const double v = c * std::sin(M_PI * (double)(n + 1) * (double)(k + 1) / (double)(N + 1));
      const size_t o = (size_t)n * s.ld + k;
      ff[o] = (float)v;
      split3(v, &bf[o], (long)count);
      if (k >= 1 && k <= N - 2) {
        fi[o - 1] = (float)v;
        split3(v, &bi[o - 1], (long)count);
      }
    }
  std::vector<float> lam(N);
  s.lam_min = 1e300;
  s.lam_max = 0.0;
  for (int k = 0; k < N; ++k) {
    const double l = (2.0 * std::cos(M_PI * (k + 1) / (double)(N + 1)) - 2.0) / (d * d);
    lam[k] = (float)l;
    s.lam_min = std::fabs(l) < s.lam_min ? std::fabs(l) : s.lam_min;
    s.lam_max = std::fabs(l) > s.lam_max ? std::fabs(l) : s.lam_max;
  }
  hipError_t e;
#define SOMAX_UP(dst, src, n)                                                            \
  if ((e = mem.alloc(&dst, n)) != hipSuccess) return e;                                 \
  if ((e = hipMemcpy(dst, src.data(), (n) * sizeof(src[0]), hipMemcpyHostToDevice)) != \
      hipSuccess)                                                                        \
    return e;
  SOMAX_UP(s.f32_full, ff, count);
  SOMAX_UP(s.f32_int, fi, count);
  SOMAX_UP(s.b_full, bf, 3 * count);
  SOMAX_UP(s.b_int, bi, 3 * count);
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
  float* dbeta_row = nullptr;
  float* lam_dev = nullptr;  // Helmholtz parameter per mode (FP32)
  float* wind_row = nullptr;
  float* homsol = nullptr;
  std::vector<double> homsol_mean;
  // workspace layout (bytes offsets), all float planes
  long wpitch = 0;  // pitch of GEMM operand planes (elements)
  long wpiece = 0;  // elements of one bf16 piece plane (maxN * wpitch)
  long splane = 0;  // Ny*Nx
  int z_split = 0;  // modes [0, z_split) use the 3-piece inverse in the mixed solve
  size_t off_OP1 = 0, off_OP2 = 0, off_QM = 0, off_P0 = 0, off_D = 0, off_Q1 = 0, off_Q2 = 0, off_PSI = 0;
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
  // row increments of beta*(y - y0), formed in FP64 (the stage kernel never
  // adds beta*y to q before differencing)
  std::vector<float> dbeta(p->Ny, 0.0f), wind(p->Ny);
  for (int j = 0; j < p->Ny; ++j) {
    if (j + 1 < p->Ny) dbeta[j] = (float)(d->beta_y_row[j + 1] - d->beta_y_row[j]);
    wind[j] = (float)d->wind_row[j];
  }
  if ((e = p->mem.alloc(&p->dbeta_row, p->Ny)) != hipSuccess ||
      (e = hipMemcpy(p->dbeta_row, dbeta.data(), p->Ny * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = p->mem.alloc(&p->wind_row, p->Ny)) != hipSuccess ||
      (e = hipMemcpy(p->wind_row, wind.data(), p->Ny * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
    set_error(std::string("somaxhip_qg_plan_create: ") + hipGetErrorString(e));
    delete p;
    return SOMAXHIP_ERR_ALLOC;
  }
  {
    std::vector<float> lf(nl);
    for (int m = 0; m < nl; ++m) lf[m] = (float)p->lambdas[m];
    if ((e = p->mem.alloc(&p->lam_dev, nl)) != hipSuccess ||
        (e = hipMemcpy(p->lam_dev, lf.data(), nl * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
      set_error(std::string("somaxhip_qg_plan_create: ") + hipGetErrorString(e));
      delete p;
      return SOMAXHIP_ERR_ALLOC;
    }
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

  // workspace layout: two GEMM operand buffers (3 bf16 pieces or 1.5 FP32
  // planes per mode), FP32 modal planes, stage states and psi.
  const long maxN = p->Nx > p->Ny ? p->Nx : p->Ny;
  p->wpitch = round_up(maxN, 64);
  p->wpiece = maxN * p->wpitch;
  p->splane = (long)p->Ny * p->Nx;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += round_up((long)bytes, 256);
    return o;
  };
  p->off_OP1 = take((size_t)nl * 3 * p->wpiece * sizeof(uint16_t));
  p->off_OP2 = take((size_t)nl * 3 * p->wpiece * sizeof(uint16_t));
  p->off_QM = take((size_t)nl * p->splane * sizeof(float));
  p->off_P0 = take((size_t)nl * p->splane * sizeof(float));
  if (p->precision == SOMAXHIP_PREC_MIXED) p->off_D = take((size_t)nl * p->splane * sizeof(float));
  p->off_Q1 = take((size_t)nl * p->splane * sizeof(float));
  p->off_Q2 = take((size_t)nl * p->splane * sizeof(float));
  p->off_PSI = take((size_t)nl * p->splane * sizeof(float));
  // Poisson-like modes (small lambda: the Laplacian's conditioning is not
  // capped) need the 3-piece inverse transform; modes come in ascending order.
  const double amp_num = p->sx.lam_max + p->sy.lam_max;
  p->z_split = 0;
  for (int m = 0; m < nl; ++m)
    if (amp_num / (p->sx.lam_min + p->sy.lam_min + p->lambdas[m]) > 1.0e4) p->z_split = m + 1;
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

const char* gemm_name(const GemmArgs& a) {
  if (a.Xf) return "dst_gemm_fp32";
  const int np = a.np_lo > a.np_hi ? a.np_lo : a.np_hi;
  return np == 1 ? "dst_gemm_bf16x1" : (np == 3 ? "dst_gemm_bf16x3" : "dst_gemm_bf16x6");
}

void gemm(const GemmArgs& a, int batch, hipStream_t st) {
  // algorithmic work: 2*R*C*K flops per plane; bytes: X read + O written once
  const double R = a.trans ? a.RB : a.RA, C = a.trans ? a.RA : a.RB;
  const double flops = 2.0 * R * C * a.K * batch;
  const double bytes = 4.0 * batch * (R * a.K + R * C);
  ProfScope ps(gemm_name(a), flops, bytes, st);
  launch_dst_gemm(a, batch, st);
}

// One DST-I layer-space solve of (lap_h - B) psi = x.
//   interior_input: the ghost ring of x is zero (q with its boundary
//   condition), so the forward transform contracts over interior points only.
//   out: psi in state layout (ld Nx); Oadd/mask as in GE_FINAL.
// Products per precision class and the bf16 pieces they read.
constexpr int NP_F32 = 0;
inline int pieces_for(int np) { return np == 1 ? 1 : (np == 3 ? 2 : 3); }

// One DST-I solve per mode of (lap_h - lambda_m) y_m = x_m, from the GEMM
// operand already in OP1 (interior rows/cols when interior_input) to the
// FP32 modal planes `out` (state layout).  np_fwd: products of the forward
// transforms; np_inv_lo / np_inv_hi: products of the inverse transforms for
// modes below / from z_split.  NP_F32 selects the FP32 kernels throughout.
void dst_solve(somaxhip_qg_plan p, bool interior_input, int np_fwd, int np_inv_lo, int np_inv_hi, float* out,
               void* ws, hipStream_t st, bool acc64_fwd = true, bool acc64_inv = true) {
  const int Nx = p->Nx, Ny = p->Ny, nl = p->nl;
  const bool f32 = np_fwd == NP_F32;
  char* OP1 = static_cast<char*>(ws) + p->off_OP1;
  char* OP2 = static_cast<char*>(ws) + p->off_OP2;
  const long pb16 = 3 * p->wpiece;      // bf16 batch stride (elements)
  const long pb32 = 3 * p->wpiece / 2;  // fp32 batch stride (floats)
  const int R0 = interior_input ? Ny - 2 : Ny;
  const int K0 = interior_input ? Nx - 2 : Nx;
  const long count_x = p->sx.rows * p->sx.ld, count_y = p->sy.rows * p->sy.ld;

  auto base = [&](GemmArgs& g, char* X, int xrows, int K, const SineSet& s, bool interior_s, int np_lo, int np_hi,
                  int zs) {
    g.ldx = p->wpitch;
    g.xrows = xrows;
    g.K = K;
    if (f32) {
      g.Xf = reinterpret_cast<const float*>(X);
      g.x_batch = pb32;
      g.Sf = interior_s ? s.f32_int : s.f32_full;
    } else {
      g.Xb = reinterpret_cast<const uint16_t*>(X);
      g.x_piece = p->wpiece;
      g.x_batch = pb16;
      g.Sb = interior_s ? s.b_int : s.b_full;
      g.s_piece = s.rows * s.ld;
    }
    g.lds = s.ld;
    g.z_split = zs;
    g.np_lo = np_lo;
    g.np_hi = np_hi;
  };
  auto out_op = [&](GemmArgs& g, char* O, int op_lo, int op_hi) {
    g.ldo = p->wpitch;
    if (f32) {
      g.Of = reinterpret_cast<float*>(O);
      g.o_batch = pb32;
    } else {
      g.Ob = reinterpret_cast<uint16_t*>(O);
      g.o_piece = p->wpiece;
      g.o_batch = pb16;
      g.op_lo = op_lo;
      g.op_hi = op_hi;
    }
  };
  const int pf = f32 ? 0 : pieces_for(np_fwd);
  const int pil = f32 ? 0 : pieces_for(np_inv_lo);
  const int pih = f32 ? 0 : pieces_for(np_inv_hi);
  const int zs = p->z_split;
  (void)count_x;
  (void)count_y;

  // F1: T[xm][y] = sum_x X[y][x] Sx[xm][x]          (S in the A slot)
  GemmArgs g{};
  base(g, OP1, R0, K0, p->sx, interior_input, np_fwd, np_fwd, 0);
  g.trans = 1;
  out_op(g, OP2, pf, pf);
  g.acc64 = acc64_fwd;
  g.RA = Nx;
  g.RB = R0;
  g.RBpad = (int)round_up(R0, 32);
  gemm(g, nl, st);

  // F2: W[xm][ym] = sum_y T[xm][y] Sy[ym][y] / (lamx + lamy - lambda)
  GemmArgs h{};
  base(h, OP2, Nx, R0, p->sy, interior_input, np_fwd, np_fwd, 0);
  h.trans = 0;
  out_op(h, OP1, pil, pih);
  h.acc64 = acc64_fwd;
  h.z_split = zs;
  h.RA = Nx;
  h.RB = Ny;
  h.RBpad = (int)round_up(Ny, 32);
  h.lamA = p->sx.lam;
  h.lamB = p->sy.lam;
  h.lam = p->lam_dev;
  gemm(h, nl, st);

  // I1: V[y][xm] = sum_ym Sy[y][ym] W[xm][ym]        (S in the A slot)
  GemmArgs u{};
  base(u, OP1, Nx, Ny, p->sy, false, np_inv_lo, np_inv_hi, zs);
  u.trans = 1;
  out_op(u, OP2, pil, pih);
  u.acc64 = acc64_inv;
  u.RA = Ny;
  u.RB = Nx;
  u.RBpad = (int)round_up(Nx, 32);
  gemm(u, nl, st);

  // I2: out[y][x] = sum_xm V[y][xm] Sx[x][xm]
  GemmArgs w{};
  base(w, OP2, Ny, Nx, p->sx, false, np_inv_lo, np_inv_hi, zs);
  w.trans = 0;
  w.Of = out;
  w.Ob = nullptr;
  w.ldo = Nx;
  w.o_batch = p->splane;
  w.op_lo = w.op_hi = 0;
  w.acc64 = acc64_inv;
  w.RA = Ny;
  w.RB = Nx;
  w.RBpad = Nx;
  gemm(w, nl, st);
}

void invert(somaxhip_qg_plan p, const float* q, float* psi, void* ws, hipStream_t st, bool apply_bc = true) {
  const int nl = p->nl;
  const bool mixed = p->precision == SOMAXHIP_PREC_MIXED;
  float* QM = wsf(ws, p->off_QM);
  float* P0 = wsf(ws, p->off_P0);
  ModalArgs ma{};
  ma.q = q;
  ma.qm = QM;
  ma.Nx = p->Nx;
  ma.Ny = p->Ny;
  ma.nl = nl;
  ma.plane = p->splane;
  ma.apply_bc = apply_bc ? 1 : 0;
  for (int a = 0; a < nl; ++a)
    for (int b = 0; b < nl; ++b) ma.cl2m[a * MAX_LAYERS + b] = p->cl2m[a * nl + b];
  ma.op = static_cast<char*>(ws) + p->off_OP1;
  ma.op_pieces = mixed ? 2 : 0;
  ma.op_ld = p->wpitch;
  ma.op_piece = p->wpiece;
  ma.op_batch = mixed ? 3 * p->wpiece : 3 * p->wpiece / 2;
  ma.op_kpad = (int)round_up(apply_bc ? p->Nx - 2 : p->Nx, 32);
  {
    ProfScope ps("to_modal", 0.0, 8.0 * nl * (double)p->splane, st);
    launch_to_modal(ma, st);
  }
  LayerArgs la{};
  la.p = P0;
  la.psi = psi;
  la.Nx = p->Nx;
  la.Ny = p->Ny;
  la.nl = nl;
  la.plane = p->splane;
  la.mask = p->zero_psi_ghost;
  for (int a = 0; a < nl; ++a)
    for (int b = 0; b < nl; ++b) la.cm2l[a * MAX_LAYERS + b] = (float)p->cm2l[a * nl + b];
  if (!mixed) {
    dst_solve(p, apply_bc, NP_F32, NP_F32, NP_F32, P0, ws, st);
  } else {
    // bf16x3 forward, x6 (Poisson-like modes) / x3 inverse; FP32 per-mode
    // Helmholtz residual; correction with bf16x3 forward and bf16 inverse.
    dst_solve(p, apply_bc, 3, 6, 3, P0, ws, st);
    float* D = wsf(ws, p->off_D);
    ResidualArgs r{};
    r.q = QM;
    r.psi = P0;
    r.r = reinterpret_cast<uint16_t*>(static_cast<char*>(ws) + p->off_OP1);
    r.op_ld = p->wpitch;
    r.op_piece = p->wpiece;
    r.op_batch = 3 * p->wpiece;
    r.op_kpad = (int)round_up(p->Nx, 32);
    r.plane = p->splane;
    r.Nx = p->Nx;
    r.Ny = p->Ny;
    r.nl = nl;
    r.idx2 = (float)(1.0 / (p->dx * p->dx));
    r.idy2 = (float)(1.0 / (p->dy * p->dy));
    for (int m = 0; m < nl; ++m) r.lam[m] = (float)p->lambdas[m];
    {
      ProfScope ps("residual", 0.0, 12.0 * nl * (double)p->splane, st);
      launch_residual(r, st);
    }
    dst_solve(p, false, 3, 1, 1, D, ws, st, true, false);
    la.d = D;
  }
  {
    ProfScope ps("to_layer", 0.0, (mixed ? 12.0 : 8.0) * nl * (double)p->splane, st);
    launch_to_layer(la, st);
  }
}

void stage(somaxhip_qg_plan p, const float* qs, const float* psi, const float* qn, float* out,
           const somaxhip_qg_params* prm, float c1, float c2, bool apply_bc, hipStream_t st) {
  StageArgs a{};
  a.qs = qs;
  a.psi = psi;
  a.qn = qn;
  a.out = out;
  a.Nx = p->Nx;
  a.Ny = p->Ny;
  a.nl = p->nl;
  a.plane = p->splane;
  a.dbeta_row = p->dbeta_row;
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
  a.c1 = c1;
  a.c2 = c2;
  a.advection = p->advection;
  a.apply_bc = apply_bc ? 1 : 0;
  // algorithmic bytes: read qs and psi, read qn when combined, write out
  const double bytes = 4.0 * p->nl * (double)p->nx * p->ny * (3.0 + (qn ? 1.0 : 0.0));
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
  stage(p, q, psi, nullptr, dq, prm, 0.0f, 1.0f, apply_bc != 0, st);
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
    stage(p, q, psi, nullptr, q1, prm, 1.0f, h, true, st);
    invert(p, q1, psi, ws, st);
    stage(p, q1, psi, q, q2, prm, 0.25f, 0.25f * h, true, st);
    invert(p, q2, psi, ws, st);
    stage(p, q2, psi, q, q, prm, 2.0f / 3.0f, 2.0f / 3.0f * h, true, st);
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
