// C ABI of the somax HIP library: batches, plans, ensembles, workspace
// layout and the QG entry points declared in include/somaxhip.h.
//
// Everything runs on batches (kernels.h): the members of one grid stacked as
// planes.  A plan is the one-member batch of a single model; an ensemble is a
// set of batches, one per grid class, over one packed state buffer.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "somaxhip.h"
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
  // folded half-length operators (even N, h = N/2): variants v = 0, 1 forward
  // S_p[i][k'] = S[2i+p][fold_phys(k')] and v = 2, 3 inverse S_p^T, each
  // frows x fld (FP32) or 3 bf16 pieces; eigenvalues lam_fold[p*h + i] = lam[2i+p]
  int h = 0;
  long fld = 0, frows = 0;
  float* f32_fold = nullptr;
  uint16_t* b_fold = nullptr;
  float* lam_fold = nullptr;
  std::vector<float> lam_host;
};

struct DeviceBlock {
  std::vector<void*> ptrs;
  DeviceBlock() = default;
  DeviceBlock(const DeviceBlock&) = delete;
  DeviceBlock& operator=(const DeviceBlock&) = delete;
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
  template <class T>
  hipError_t upload(T** out, const std::vector<T>& v) {
    hipError_t e = alloc(out, v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
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
  if ((e = mem.upload(&s.f32_full, ff)) != hipSuccess) return e;
  if ((e = mem.upload(&s.f32_int, fi)) != hipSuccess) return e;
  if ((e = mem.upload(&s.b_full, bf)) != hipSuccess) return e;
  if ((e = mem.upload(&s.b_int, bi)) != hipSuccess) return e;
  s.lam_host = lam;
  return mem.upload(&s.lam, lam);
}

// Folded operators of an even-length axis (S[N-1-n][a] = (-1)^a S[n][a]).
static hipError_t build_fold(DeviceBlock& mem, SineSet& s) {
  const int N = s.N, h = N / 2;
  s.h = h;
  s.fld = round_up(h, 32);
  s.frows = round_up(h, 128);
  const size_t var = (size_t)s.frows * s.fld;
  std::vector<float> ff(4 * var, 0.0f), lf(2 * (size_t)h);
  std::vector<uint16_t> bf(4 * 3 * var, 0);
  const double c = std::sqrt(2.0 / (N + 1));
  auto S = [&](int a, int n) { return c * std::sin(M_PI * (double)(a + 1) * (double)(n + 1) / (double)(N + 1)); };
  for (int p = 0; p < 2; ++p)
    for (int i = 0; i < h; ++i)
      for (int k = 0; k < h; ++k) {
        const double v = S(2 * i + p, fold_phys(k, h));
        const size_t fwd = (size_t)p * var + (size_t)i * s.fld + k;       // rows = outputs i
        const size_t inv = (size_t)(2 + p) * var + (size_t)k * s.fld + i;  // rows = outputs k'
        ff[fwd] = (float)v;
        ff[inv] = (float)v;
        split3(v, &bf[(size_t)p * 3 * var + (size_t)i * s.fld + k], (long)var);
        split3(v, &bf[(size_t)(2 + p) * 3 * var + (size_t)k * s.fld + i], (long)var);
      }
  for (int p = 0; p < 2; ++p)
    for (int i = 0; i < h; ++i) lf[(size_t)p * h + i] = s.lam_host[2 * i + p];
  hipError_t e;
  if ((e = mem.upload(&s.f32_fold, ff)) != hipSuccess) return e;
  if ((e = mem.upload(&s.b_fold, bf)) != hipSuccess) return e;
  return mem.upload(&s.lam_fold, lf);
}

// Folded transforms halve the GEMM work; below h = 128 the four quarter
// blocks cost more tiles than the plain transform.  SOMAXHIP_FOLD=0 / 1
// turns them off / on for every even grid.
static bool use_fold(int Nx, int Ny) {
  if (Nx % 2 || Ny % 2) return false;
  const char* env = std::getenv("SOMAXHIP_FOLD");
  if (env && env[0] == '0') return false;
  if (env && env[0] == '1') return Nx >= 4 && Ny >= 4;
  return Nx / 2 >= 128 && Ny / 2 >= 128;
}

// The members of one grid, stacked as planes, with their device tables and
// the layout of the batch's workspace.
struct Batch {
  int nx = 0, ny = 0, Nx = 0, Ny = 0;
  double dx = 0.0, dy = 0.0;
  int M = 0, P = 0;
  int advection = 0, precision = 0;
  int fold = 0;           // folded transforms (4 parity blocks of (Ny/2) x (Nx/2) per plane)
  long fplane = 0;        // folded: (Ny/2) * (Nx/2)
  DeviceBlock mem;
  SineSet sx, sy;
  Tables t{};
  float* dbeta_rows = nullptr;
  float* wind_rows = nullptr;
  std::vector<int> first, nl;  // host copies of the member tables
  // workspace layout (byte offsets), all float planes
  long wpitch = 0;  // pitch of GEMM operand planes (elements)
  long wpiece = 0;  // elements of one bf16 piece plane (maxN * wpitch)
  long splane = 0;  // Ny*Nx
  size_t off_OP1 = 0, off_OP2 = 0, off_QM = 0, off_P0 = 0, off_D = 0, off_Q1 = 0, off_Q2 = 0, off_PSI = 0;
  size_t off_DIAG = 0;  // FP64 partial sums of the diagnostics and the mass constraint
  size_t off_PWALL = 0;  // per-plane wall psi
  size_t ws_bytes = 0;
  int stage_rows = 16;  // Arakawa stage tile height (8 or 16), chosen per class
  // mass correction: homogeneous solutions (P planes) and their means (0: none)
  int any_mass = 0;
  float* homsol = nullptr;
  double* hmean = nullptr;
};

// Checks one model description; empty string when valid.
static std::string check_desc(const somaxhip_qg_desc* d, bool allow_f64) {
  if (!d) return "null description";
  if (d->nx < 2 || d->ny < 2 || d->nl < 1 || d->nl > MAX_LAYERS || !(d->dx > 0) || !(d->dy > 0))
    return "need nx, ny >= 2, 1 <= nl <= 8, dx, dy > 0";
  if ((long)(d->nx + 2) * (d->ny + 2) > 0x7fffffffL) return "a plane of (nx + 2) * (ny + 2) cells must be below 2^31";
  if (!d->cl2m || !d->cm2l || !d->lambdas || !d->beta_y_row || !d->wind_row || !d->wind_scale)
    return "a required table pointer is null";
  if (d->advection != SOMAXHIP_ADV_ARAKAWA && d->advection != SOMAXHIP_ADV_WENO) return "unknown advection";
  if (d->precision == SOMAXHIP_PREC_FP64 && !allow_f64)
    return "ensembles compute in FP32 / BF16; SOMAXHIP_PREC_FP64 is a single-model form";
  if (d->precision != SOMAXHIP_PREC_FP32 && d->precision != SOMAXHIP_PREC_MIXED &&
      d->precision != SOMAXHIP_PREC_BF16X6 && d->precision != SOMAXHIP_PREC_FP64)
    return "unknown precision";
  if (d->mass_correction && (!d->homsol || !d->homsol_mean))
    return "mass_correction needs homsol and homsol_mean";
  return "";
}

static bool same_class(const somaxhip_qg_desc* a, const somaxhip_qg_desc* b) {
  return a->nx == b->nx && a->ny == b->ny && a->dx == b->dx && a->dy == b->dy && a->advection == b->advection &&
         a->precision == b->precision;
}

// Builds batch `b` from M descriptions of one class (validated by the caller).
static hipError_t batch_build(Batch& b, const std::vector<const somaxhip_qg_desc*>& ds) {
  const somaxhip_qg_desc* d0 = ds[0];
  b.nx = d0->nx;
  b.ny = d0->ny;
  b.Nx = d0->nx + 2;
  b.Ny = d0->ny + 2;
  b.dx = d0->dx;
  b.dy = d0->dy;
  b.advection = d0->advection;
  b.precision = d0->precision;
  b.M = (int)ds.size();
  hipError_t e;
  if ((e = build_sines(b.mem, b.sx, b.Nx, b.dx)) != hipSuccess) return e;
  if ((e = build_sines(b.mem, b.sy, b.Ny, b.dy)) != hipSuccess) return e;
  b.fold = use_fold(b.Nx, b.Ny) ? 1 : 0;
  if (b.fold) {
    if ((e = build_fold(b.mem, b.sx)) != hipSuccess) return e;
    if ((e = build_fold(b.mem, b.sy)) != hipSuccess) return e;
    b.fplane = (long)b.sx.h * b.sy.h;
  }

  // Poisson-like modes (small lambda: the Laplacian's conditioning is not
  // capped) need the 3-piece inverse transform in the mixed solve.
  const double amp_num = b.sx.lam_max + b.sy.lam_max;
  const double amp_den0 = b.sx.lam_min + b.sy.lam_min;

  std::vector<int> mfirst, mnl, member, layer;
  std::vector<double> cl2m((size_t)b.M * 64, 0.0);
  std::vector<float> cm2l((size_t)b.M * 64, 0.0f), lam, wscale;
  std::vector<unsigned char> mask, hp;
  std::vector<float> dbeta((size_t)b.M * b.Ny, 0.0f), wind((size_t)b.M * b.Ny, 0.0f);
  int P = 0;
  for (int m = 0; m < b.M; ++m) {
    const somaxhip_qg_desc* d = ds[m];
    const int nl = d->nl;
    mfirst.push_back(P);
    mnl.push_back(nl);
    mask.push_back(d->zero_psi_ghost ? 1 : 0);
    for (int i = 0; i < nl; ++i)
      for (int k = 0; k < nl; ++k) {
        cl2m[(size_t)m * 64 + i * MAX_LAYERS + k] = d->cl2m[i * nl + k];
        cm2l[(size_t)m * 64 + i * MAX_LAYERS + k] = (float)d->cm2l[i * nl + k];
      }
    for (int l = 0; l < nl; ++l) {
      member.push_back(m);
      layer.push_back(l);
      lam.push_back((float)d->lambdas[l]);
      const double den = amp_den0 + d->lambdas[l];
      hp.push_back(!(den > 0.0) || amp_num / den > 1.0e4 ? 1 : 0);
      wscale.push_back((float)d->wind_scale[l]);
    }
    // row increments of beta*(y - y0), formed in FP64 (the stage kernel never
    // adds beta*y to q before differencing)
    for (int j = 0; j < b.Ny; ++j) {
      if (j + 1 < b.Ny) dbeta[(size_t)m * b.Ny + j] = (float)(d->beta_y_row[j + 1] - d->beta_y_row[j]);
      wind[(size_t)m * b.Ny + j] = (float)d->wind_row[j];
    }
    P += nl;
  }
  b.P = P;
  b.first = mfirst;
  b.nl = mnl;
  int *dfirst, *dnl, *dmember, *dlayer;
  double* dcl2m;
  float *dcm2l, *dlam, *dws;
  unsigned char *dmask, *dhp;
  if ((e = b.mem.upload(&dfirst, mfirst)) != hipSuccess || (e = b.mem.upload(&dnl, mnl)) != hipSuccess ||
      (e = b.mem.upload(&dcl2m, cl2m)) != hipSuccess || (e = b.mem.upload(&dcm2l, cm2l)) != hipSuccess ||
      (e = b.mem.upload(&dmask, mask)) != hipSuccess || (e = b.mem.upload(&dmember, member)) != hipSuccess ||
      (e = b.mem.upload(&dlayer, layer)) != hipSuccess || (e = b.mem.upload(&dlam, lam)) != hipSuccess ||
      (e = b.mem.upload(&dhp, hp)) != hipSuccess || (e = b.mem.upload(&dws, wscale)) != hipSuccess ||
      (e = b.mem.upload(&b.dbeta_rows, dbeta)) != hipSuccess || (e = b.mem.upload(&b.wind_rows, wind)) != hipSuccess)
    return e;
  b.t = Tables{dfirst, dnl, dcl2m, dcm2l, dmask, dmember, dlayer, dlam, dhp, dws, nullptr};

  // mass correction: homogeneous solution and its mean per plane
  for (const somaxhip_qg_desc* d : ds) b.any_mass |= d->mass_correction ? 1 : 0;
  if (b.any_mass) {
    const size_t plane = (size_t)b.Ny * b.Nx;
    std::vector<float> hs((size_t)P * plane, 0.0f);
    std::vector<double> hm(P, 0.0);
    for (int m = 0; m < b.M; ++m) {
      const somaxhip_qg_desc* d = ds[m];
      if (!d->mass_correction) continue;
      std::copy(d->homsol, d->homsol + (size_t)d->nl * plane, hs.begin() + (size_t)mfirst[m] * plane);
      for (int l = 0; l < d->nl; ++l) hm[mfirst[m] + l] = d->homsol_mean[l];
    }
    if ((e = b.mem.upload(&b.homsol, hs)) != hipSuccess || (e = b.mem.upload(&b.hmean, hm)) != hipSuccess) return e;
  }

  // workspace: two GEMM operand buffers (3 bf16 pieces or 1.5 FP32 planes
  // per plane), FP32 modal planes, stage states and psi.
  // (folded: four blocks per plane, each at most (maxN/2) x pitch)
  const long maxN = b.Nx > b.Ny ? b.Nx : b.Ny;
  const long rowsN = b.fold ? maxN / 2 : maxN;
  b.wpitch = round_up(rowsN, 64);
  b.wpiece = rowsN * b.wpitch;
  b.splane = (long)b.Ny * b.Nx;
  const int opl = b.fold ? 4 * P : P;  // GEMM operand planes
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += round_up((long)bytes, 256);
    return o;
  };
  b.off_OP1 = take((size_t)opl * 3 * b.wpiece * sizeof(uint16_t));
  b.off_OP2 = take((size_t)opl * 3 * b.wpiece * sizeof(uint16_t));
  b.off_QM = take((size_t)P * b.splane * sizeof(float));
  b.off_P0 = take((size_t)P * b.splane * sizeof(float));
  if (b.precision == SOMAXHIP_PREC_MIXED) b.off_D = take((size_t)P * b.splane * sizeof(float));
  b.off_Q1 = take((size_t)P * b.splane * sizeof(float));
  b.off_Q2 = take((size_t)P * b.splane * sizeof(float));
  b.off_PSI = take((size_t)P * b.splane * sizeof(float));
  b.off_DIAG = take(qg_diag_partial_doubles(b.Nx, b.Ny, P) * sizeof(double));
  b.off_PWALL = take((size_t)P * sizeof(float));
  b.ws_bytes = off;
  return prepare_dst_gemm();
}

static MemberParams to_member_params(const somaxhip_qg_params& p) {
  return MemberParams{(float)p.nu, (float)p.nu4, (float)p.kappa, (float)p.tau0};
}

}  // namespace somax

using namespace somax;

struct somaxhip_qg_plan_s {
  Batch b;               // FP32 / BF16 plans
  Qg64* f64 = nullptr;   // FP64 plans (b stays empty)
  ~somaxhip_qg_plan_s() {
    if (f64) qg64_destroy(f64);
  }
};

namespace somax {
int qg_plan_shape(somaxhip_qg_plan p, int* planes, int* Ny, int* Nx) {
  if (!p) return SOMAXHIP_ERR_ARG;
  if (p->f64) {
    set_error("the plan computes in FP64; this call takes an FP32 / BF16 plan");
    return SOMAXHIP_ERR_ARG;
  }
  *planes = p->b.P;
  *Ny = p->b.Ny;
  *Nx = p->b.Nx;
  return SOMAXHIP_OK;
}
const Qg64* qg_plan_f64(somaxhip_qg_plan p) { return p ? p->f64 : nullptr; }
}  // namespace somax

struct somaxhip_ens_s {
  std::vector<std::unique_ptr<Batch>> cls;  // grid classes, largest first
  std::vector<int> mclass;                  // input member -> class
  std::vector<long> moff;                   // input member -> float offset in the packed state
  std::vector<size_t> cstate;               // class -> float offset of its planes
  std::vector<size_t> cws;                  // class -> byte offset of its workspace
  std::vector<std::vector<int>> cmembers;   // class -> input member of each local member
  std::vector<MemberParams*> cparams;       // class -> device per-member parameters
  std::vector<std::vector<MemberParams>> hparams;
  std::vector<int> cbiharm;
  size_t state_floats = 0, ws_bytes = 0;
  int flags = 0;
  // concurrency and graph state
  hipStream_t main = nullptr;
  std::vector<hipStream_t> side;
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  hipGraphExec_t exec = nullptr;
  const void* g_q = nullptr;
  const void* g_ws = nullptr;
  double g_dt = 0.0;
  bool stale = false;
  bool stepped = false;  // launch shapes are fixed from the first step on
  ~somaxhip_ens_s() {
    if (main) (void)hipStreamSynchronize(main);
    if (exec) (void)hipGraphExecDestroy(exec);
    for (hipStream_t s : side) (void)hipStreamDestroy(s);
    if (main) (void)hipStreamDestroy(main);
    for (hipEvent_t e : ev_join) (void)hipEventDestroy(e);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
  }
};

// ---------------------------------------------------------------------------
// Step without a caller workspace: the workspace is `scratch`, a device array
// of SOMAXHIP_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds
// from module load until the process ends.  In host code `scratch` names a
// host-side stand-in; its device address on the current device comes from
// hipGetSymbolAddress once per device and is the only form passed to kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[SOMAXHIP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(void** out, const char* who) {
  int dev = 0;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    set_error(std::string(who) + ": no current device (" + hipGetErrorString(e) + ")");
    return SOMAXHIP_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t es = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (es != hipSuccess || !ptr) {
      set_error(std::string(who) + ": hipGetSymbolAddress(scratch) failed: " + hipGetErrorString(es));
      return SOMAXHIP_ERR_HIP;
    }
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return SOMAXHIP_OK;
}

}  // namespace

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

int somaxhip_version(void) { return 200; }

int somaxhip_init(void) {
  int dev = 0;
  SOMAX_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SOMAX_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("somaxhip is built for gfx950; current device is ") + prop.gcnArchName);
    return SOMAXHIP_ERR_DEVICE;
  }
  void* ptr = nullptr;  // this device's address of the no-allocation step's scratch array
  return scratch_address(&ptr, "somaxhip_init");
}

int somaxhip_qg_plan_create(const somaxhip_qg_desc* d, somaxhip_qg_plan* out) {
  if (!d || !out) {
    set_error("somaxhip_qg_plan_create: null argument");
    return SOMAXHIP_ERR_ARG;
  }
  *out = nullptr;
  const std::string bad = check_desc(d, true);
  if (!bad.empty()) {
    set_error("somaxhip_qg_plan_create: " + bad);
    return SOMAXHIP_ERR_ARG;
  }
  somaxhip_qg_plan p = new (std::nothrow) somaxhip_qg_plan_s();
  if (!p) {
    set_error("somaxhip_qg_plan_create: out of host memory");
    return SOMAXHIP_ERR_ALLOC;
  }
  if (d->precision == SOMAXHIP_PREC_FP64) {
    std::string err;
    p->f64 = qg64_create(d, &err);
    if (!p->f64) {
      set_error("somaxhip_qg_plan_create: " + err);
      delete p;
      return SOMAXHIP_ERR_ALLOC;
    }
    *out = p;
    return SOMAXHIP_OK;
  }
  const hipError_t e = batch_build(p->b, {d});
  if (e != hipSuccess) {
    set_error(std::string("somaxhip_qg_plan_create: ") + hipGetErrorString(e));
    delete p;
    return SOMAXHIP_ERR_ALLOC;
  }
  *out = p;
  return SOMAXHIP_OK;
}

int somaxhip_qg_plan_destroy(somaxhip_qg_plan p) {
  delete p;
  return SOMAXHIP_OK;
}

size_t somaxhip_qg_workspace_bytes(somaxhip_qg_plan p) {
  if (!p) return 0;
  return p->f64 ? qg64_workspace_bytes(p->f64) : p->b.ws_bytes;
}

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

// Products of a GEMM class: NP_F32 selects the FP32 kernels; 1, 3, 6 bf16.
constexpr int NP_F32 = 0;

// The four transforms of one solve (F1, F2, I1, I2) with their batch sizes,
// recorded instead of launched when a merged launch will carry them.
struct SolveGemms {
  GemmArgs g[4];
  int batch[4];
};
// Launch transform k of a solve, or record it into `rec`.
inline void emit_gemm(const GemmArgs& a, int batch, int k, SolveGemms* rec, hipStream_t st) {
  if (rec) {
    rec->g[k] = a;
    rec->batch[k] = batch;
    return;
  }
  gemm(a, batch, st);
}
inline int pieces_for(int np) { return np == 1 ? 1 : (np == 3 ? 2 : 3); }

// Folded solve: the four parity blocks of every plane as 4P half-size planes
// (plane b * P + m).  In: OP1 holds the fold sums (interior input: the ghost
// pair, last in the folded order, is zero and skipped).  Out: `out` receives
// the parity components of the solution, 4P planes of hy x hx.
void dst_solve_folded(const Batch& p, bool interior_input, int np_fwd, int np_inv_lo, int np_inv_hi, float* out,
                      void* ws, hipStream_t st, bool acc64_fwd, bool acc64_inv, SolveGemms* rec) {
  const int hx = p.sx.h, hy = p.sy.h, P = p.P;
  const bool f32 = np_fwd == NP_F32;
  char* OP1 = static_cast<char*>(ws) + p.off_OP1;
  char* OP2 = static_cast<char*>(ws) + p.off_OP2;
  const long pb16 = 3 * p.wpiece;
  const long pb32 = 3 * p.wpiece / 2;
  const int R0 = interior_input ? hy - 1 : hy;  // F1 data rows = F2 k extent
  const int K0 = interior_input ? hx - 1 : hx;  // F1 k extent

  // variant v of an axis: 0/1 forward parity, 2/3 inverse parity
  auto base = [&](GemmArgs& g, char* X, int xrows, int K, const SineSet& s, int inverse, int s_par, int np_lo,
                  int np_hi) {
    const long var = s.frows * s.fld;
    g.ldx = p.wpitch;
    g.xrows = xrows;
    g.K = K;
    if (f32) {
      g.Xf = reinterpret_cast<const float*>(X);
      g.x_batch = pb32;
      g.Sf = s.f32_fold + (inverse ? 2 * var : 0);
      g.s_var = var;
    } else {
      g.Xb = reinterpret_cast<const uint16_t*>(X);
      g.x_piece = p.wpiece;
      g.x_batch = pb16;
      g.Sb = s.b_fold + (inverse ? 2 * 3 * var : 0);
      g.s_piece = var;
      g.s_var = 3 * var;
    }
    g.lds = s.fld;
    g.hp = p.t.hp;
    g.np_lo = np_lo;
    g.np_hi = np_hi;
    g.fold_P = P;
    g.s_par = s_par;
  };
  auto out_op = [&](GemmArgs& g, char* O, int op_lo, int op_hi) {
    g.ldo = p.wpitch;
    if (f32) {
      g.Of = reinterpret_cast<float*>(O);
      g.o_batch = pb32;
    } else {
      g.Ob = reinterpret_cast<uint16_t*>(O);
      g.o_piece = p.wpiece;
      g.o_batch = pb16;
      g.op_lo = op_lo;
      g.op_hi = op_hi;
    }
  };
  const int pf = f32 ? 0 : pieces_for(np_fwd);
  const int pil = f32 ? 0 : pieces_for(np_inv_lo);
  const int pih = f32 ? 0 : pieces_for(np_inv_hi);

  // F1: T[ix][y'] = sum_x' Sx_q[ix][x'] X[y'][x']
  GemmArgs g{};
  base(g, OP1, R0, K0, p.sx, 0, 0, np_fwd, np_fwd);
  g.trans = 1;
  out_op(g, OP2, pf, pf);
  g.acc64 = acc64_fwd;
  g.RA = hx;
  g.RB = R0;
  g.RBpad = (int)round_up(R0, 32);
  emit_gemm(g, 4 * P, 0, rec, st);

  // F2: W[ix][iy] = sum_y' T[ix][y'] Sy_p[iy][y'] / (lamx_q[ix] + lamy_p[iy] - lambda)
  GemmArgs h{};
  base(h, OP2, hx, R0, p.sy, 0, 1, np_fwd, np_fwd);
  h.trans = 0;
  out_op(h, OP1, pil, pih);
  h.acc64 = acc64_fwd;
  h.RA = hx;
  h.RB = hy;
  h.RBpad = (int)round_up(hy, 32);
  h.lamA = p.sx.lam_fold;
  h.lamB = p.sy.lam_fold;
  h.lam_var = hx;
  h.lamB_var = hy;
  h.lam = p.t.lam;
  emit_gemm(h, 4 * P, 1, rec, st);

  // I1: V[y'][ix] = sum_iy Sy_p^T[y'][iy] W[ix][iy]
  GemmArgs u{};
  base(u, OP1, hx, hy, p.sy, 1, 1, np_inv_lo, np_inv_hi);
  u.trans = 1;
  out_op(u, OP2, pil, pih);
  u.acc64 = acc64_inv;
  u.RA = hy;
  u.RB = hx;
  u.RBpad = (int)round_up(hx, 32);
  emit_gemm(u, 4 * P, 2, rec, st);

  // I2: out[y'][x'] = sum_ix V[y'][ix] Sx_q^T[x'][ix]
  GemmArgs w{};
  base(w, OP2, hy, hx, p.sx, 1, 0, np_inv_lo, np_inv_hi);
  w.trans = 0;
  w.Of = out;
  w.Ob = nullptr;
  w.ldo = hx;
  w.o_batch = p.fplane;
  w.op_lo = w.op_hi = 0;
  w.acc64 = acc64_inv;
  w.RA = hy;
  w.RB = hx;
  w.RBpad = hx;
  emit_gemm(w, 4 * P, 3, rec, st);
}

// One DST-I solve per plane of (lap_h - lambda_z) y_z = x_z, from the GEMM
// operand already in OP1 (interior rows/cols when interior_input: the ghost
// ring of x is zero, so the forward transform contracts over interior points
// only) to the FP32 modal planes `out` (state layout; folded batches: parity
// components).  np_fwd: products of the forward transforms; np_inv_lo /
// np_inv_hi: products of the inverse transforms of Poisson-like / other planes.
void dst_solve(const Batch& p, bool interior_input, int np_fwd, int np_inv_lo, int np_inv_hi, float* out,
               void* ws, hipStream_t st, bool acc64_fwd = true, bool acc64_inv = true, SolveGemms* rec = nullptr) {
  if (p.fold) {
    dst_solve_folded(p, interior_input, np_fwd, np_inv_lo, np_inv_hi, out, ws, st, acc64_fwd, acc64_inv, rec);
    return;
  }
  const int Nx = p.Nx, Ny = p.Ny, P = p.P;
  const bool f32 = np_fwd == NP_F32;
  char* OP1 = static_cast<char*>(ws) + p.off_OP1;
  char* OP2 = static_cast<char*>(ws) + p.off_OP2;
  const long pb16 = 3 * p.wpiece;      // bf16 batch stride (elements)
  const long pb32 = 3 * p.wpiece / 2;  // fp32 batch stride (floats)
  const int R0 = interior_input ? Ny - 2 : Ny;
  const int K0 = interior_input ? Nx - 2 : Nx;

  auto base = [&](GemmArgs& g, char* X, int xrows, int K, const SineSet& s, bool interior_s, int np_lo,
                  int np_hi) {
    g.ldx = p.wpitch;
    g.xrows = xrows;
    g.K = K;
    if (f32) {
      g.Xf = reinterpret_cast<const float*>(X);
      g.x_batch = pb32;
      g.Sf = interior_s ? s.f32_int : s.f32_full;
    } else {
      g.Xb = reinterpret_cast<const uint16_t*>(X);
      g.x_piece = p.wpiece;
      g.x_batch = pb16;
      g.Sb = interior_s ? s.b_int : s.b_full;
      g.s_piece = s.rows * s.ld;
    }
    g.lds = s.ld;
    g.hp = p.t.hp;
    g.np_lo = np_lo;
    g.np_hi = np_hi;
  };
  auto out_op = [&](GemmArgs& g, char* O, int op_lo, int op_hi) {
    g.ldo = p.wpitch;
    if (f32) {
      g.Of = reinterpret_cast<float*>(O);
      g.o_batch = pb32;
    } else {
      g.Ob = reinterpret_cast<uint16_t*>(O);
      g.o_piece = p.wpiece;
      g.o_batch = pb16;
      g.op_lo = op_lo;
      g.op_hi = op_hi;
    }
  };
  const int pf = f32 ? 0 : pieces_for(np_fwd);
  const int pil = f32 ? 0 : pieces_for(np_inv_lo);
  const int pih = f32 ? 0 : pieces_for(np_inv_hi);

  // F1: T[xm][y] = sum_x X[y][x] Sx[xm][x]          (S in the A slot)
  GemmArgs g{};
  base(g, OP1, R0, K0, p.sx, interior_input, np_fwd, np_fwd);
  g.trans = 1;
  out_op(g, OP2, pf, pf);
  g.acc64 = acc64_fwd;
  g.RA = Nx;
  g.RB = R0;
  g.RBpad = (int)round_up(R0, 32);
  emit_gemm(g, P, 0, rec, st);

  // F2: W[xm][ym] = sum_y T[xm][y] Sy[ym][y] / (lamx + lamy - lambda)
  GemmArgs h{};
  base(h, OP2, Nx, R0, p.sy, interior_input, np_fwd, np_fwd);
  h.trans = 0;
  out_op(h, OP1, pil, pih);
  h.acc64 = acc64_fwd;
  h.RA = Nx;
  h.RB = Ny;
  h.RBpad = (int)round_up(Ny, 32);
  h.lamA = p.sx.lam;
  h.lamB = p.sy.lam;
  h.lam = p.t.lam;
  emit_gemm(h, P, 1, rec, st);

  // I1: V[y][xm] = sum_ym Sy[y][ym] W[xm][ym]        (S in the A slot)
  GemmArgs u{};
  base(u, OP1, Nx, Ny, p.sy, false, np_inv_lo, np_inv_hi);
  u.trans = 1;
  out_op(u, OP2, pil, pih);
  u.acc64 = acc64_inv;
  u.RA = Ny;
  u.RB = Nx;
  u.RBpad = (int)round_up(Nx, 32);
  emit_gemm(u, P, 2, rec, st);

  // I2: out[y][x] = sum_xm V[y][xm] Sx[x][xm]
  GemmArgs w{};
  base(w, OP2, Ny, Nx, p.sx, false, np_inv_lo, np_inv_hi);
  w.trans = 0;
  w.Of = out;
  w.Ob = nullptr;
  w.ldo = Nx;
  w.o_batch = p.splane;
  w.op_lo = w.op_hi = 0;
  w.acc64 = acc64_inv;
  w.RA = Ny;
  w.RB = Nx;
  w.RBpad = Nx;
  emit_gemm(w, P, 3, rec, st);
}

// Arguments of the inversion's kernels for one batch.
ModalArgs modal_args(const Batch& p, const float* q, void* ws, bool apply_bc) {
  const bool mixed = p.precision == SOMAXHIP_PREC_MIXED;
  const bool x6 = p.precision == SOMAXHIP_PREC_BF16X6;
  ModalArgs ma{};
  ma.q = q;
  ma.qm = wsf(ws, p.off_QM);
  ma.Nx = p.Nx;
  ma.Ny = p.Ny;
  ma.M = p.M;
  ma.plane = p.splane;
  ma.apply_bc = apply_bc ? 1 : 0;
  ma.t = p.t;
  ma.op = static_cast<char*>(ws) + p.off_OP1;
  ma.op_pieces = mixed ? 2 : (x6 ? 3 : 0);
  ma.op_ld = p.wpitch;
  ma.op_piece = p.wpiece;
  ma.op_batch = (mixed || x6) ? 3 * p.wpiece : 3 * p.wpiece / 2;
  ma.op_kpad = (int)round_up(apply_bc ? p.Nx - 2 : p.Nx, 32);
  ma.nlmax = *std::max_element(p.nl.begin(), p.nl.end());
  if (p.fold) {
    ma.fold = 1;
    ma.P = p.P;
    ma.op_kpad = (int)round_up(apply_bc ? p.sx.h - 1 : p.sx.h, 32);
  }
  return ma;
}

ResidualArgs residual_args(const Batch& p, void* ws) {
  ResidualArgs r{};
  r.q = wsf(ws, p.off_QM);
  r.psi = wsf(ws, p.off_P0);
  r.r = reinterpret_cast<uint16_t*>(static_cast<char*>(ws) + p.off_OP1);
  r.op_ld = p.wpitch;
  r.op_piece = p.wpiece;
  r.op_batch = 3 * p.wpiece;
  r.op_kpad = (int)round_up(p.fold ? p.sx.h : p.Nx, 32);
  r.fold = p.fold;
  r.plane = p.splane;
  r.Nx = p.Nx;
  r.Ny = p.Ny;
  r.P = p.P;
  r.idx2 = (float)(1.0 / (p.dx * p.dx));
  r.idy2 = (float)(1.0 / (p.dy * p.dy));
  r.lam = p.t.lam;
  return r;
}

// to_layer arguments; the correction plane when `with_correction`
LayerArgs layer_args(const Batch& p, float* psi, void* ws, bool with_correction) {
  LayerArgs la{};
  la.pm = wsf(ws, p.off_P0);
  la.d = with_correction ? wsf(ws, p.off_D) : nullptr;
  la.psi = psi;
  la.Nx = p.Nx;
  la.Ny = p.Ny;
  la.M = p.M;
  la.plane = p.splane;
  la.t = p.t;
  la.fold = p.fold;
  la.P = p.P;
  la.nlmax = *std::max_element(p.nl.begin(), p.nl.end());
  return la;
}

// The mass constraint's plane sums of the modal psi (first guess plus the
// correction), FP64 partials in the diagnostics area; folded: the array sum
// is 4x the sum of the even-even block (planes 0..P-1).
PlaneSumArgs mass_sums_args(const Batch& p, const LayerArgs& la, void* ws) {
  double* partial = reinterpret_cast<double*>(static_cast<char*>(ws) + p.off_DIAG);
  return plane_sums_args(la.pm, la.d, p.fold ? p.sx.h : p.Nx, p.fold ? p.sy.h : p.Ny, p.P,
                         p.fold ? p.fplane : p.splane, partial);
}
// to_layer forms alpha = -mean / mean(homsol) from the partials and the wall psi
void mass_layer_args(const Batch& p, LayerArgs& la, const PlaneSumArgs& ps, void* ws) {
  la.nblk = ps.nblk;
  la.homsol = p.homsol;
  la.partial = ps.partial;
  la.hmean = p.hmean;
  la.inv_count = (p.fold ? 4.0 : 1.0) / (double)p.splane;
  la.psi_wall = wsf(ws, p.off_PWALL);
}

void invert(const Batch& p, const float* q, float* psi, void* ws, hipStream_t st, bool apply_bc = true) {
  const bool mixed = p.precision == SOMAXHIP_PREC_MIXED;
  const bool x6 = p.precision == SOMAXHIP_PREC_BF16X6;
  float* P0 = wsf(ws, p.off_P0);
  {
    const ModalArgs ma = modal_args(p, q, ws, apply_bc);
    ProfScope ps("to_modal", 0.0, 8.0 * p.P * (double)p.splane, st);
    launch_to_modal(ma, st);
  }
  if (x6) {
    // three-piece operands, six products and FP64 flushes in every transform:
    // about 2^-24 per product, accurate without a refinement pass
    dst_solve(p, apply_bc, 6, 6, 6, P0, ws, st);
  } else if (!mixed) {
    dst_solve(p, apply_bc, NP_F32, NP_F32, NP_F32, P0, ws, st);
  } else {
    // bf16x3 forward, x6 (Poisson-like modes) / x3 inverse; FP32 per-mode
    // Helmholtz residual; correction with bf16x3 forward and bf16 inverse.
    dst_solve(p, apply_bc, 3, 6, 3, P0, ws, st);
    {
      const ResidualArgs r = residual_args(p, ws);
      ProfScope ps("residual", 0.0, 12.0 * p.P * (double)p.splane, st);
      launch_residual(r, st);
    }
    dst_solve(p, false, 3, 1, 1, wsf(ws, p.off_D), ws, st, true, false);
  }
  LayerArgs la = layer_args(p, psi, ws, mixed);
  if (p.any_mass) {
    // FP64 domain mean of every modal streamfunction, alpha = -mean / mean(homsol)
    const PlaneSumArgs sums = mass_sums_args(p, la, ws);
    {
      ProfScope ps("mass_constraint", 0.0, (mixed ? 8.0 : 4.0) * p.P * (double)p.splane, st);
      launch_plane_sums(sums.a, sums.b, p.fold ? p.sx.h : p.Nx, p.fold ? p.sy.h : p.Ny, p.P, sums.plane,
                        sums.partial, st);
    }
    mass_layer_args(p, la, sums, ws);
  }
  {
    ProfScope ps("to_layer", 0.0, (mixed ? 12.0 : 8.0) * p.P * (double)p.splane, st);
    launch_to_layer(la, st);
  }
}

// One batch of a merged launch group: its state, streamfunction and workspace.
struct GroupItem {
  const Batch* b;
  const float* q;
  float* psi;
  void* ws;
};

// invert() for the batches of a group (same precision): every kernel phase is
// one launch over all of them; each batch's arithmetic is that of invert().
void invert_group(const std::vector<GroupItem>& it, hipStream_t st) {
  const int n = (int)it.size();
  const int prec = it[0].b->precision;
  const bool mixed = prec == SOMAXHIP_PREC_MIXED;
  const bool x6 = prec == SOMAXHIP_PREC_BF16X6;
  double bytes = 0.0;
  for (const GroupItem& g : it) bytes += (double)g.b->P * g.b->splane;
  {
    ModalArgs ma[MAX_GROUP];
    for (int i = 0; i < n; ++i) ma[i] = modal_args(*it[i].b, it[i].q, it[i].ws, true);
    ProfScope ps("to_modal", 0.0, 8.0 * bytes, st);
    launch_to_modal_group(ma, n, st);
  }
  auto solves = [&](bool interior, int npf, int npl, int nph, bool correction, bool acc64f, bool acc64i) {
    SolveGemms rec[MAX_GROUP];
    for (int i = 0; i < n; ++i) {
      const Batch& p = *it[i].b;
      dst_solve(p, interior, npf, npl, nph, wsf(it[i].ws, correction ? p.off_D : p.off_P0), it[i].ws, st, acc64f,
                acc64i, &rec[i]);
    }
    for (int k = 0; k < 4; ++k) {
      GemmArgs g[MAX_GROUP];
      int batch[MAX_GROUP];
      double flops = 0.0, gbytes = 0.0;
      for (int i = 0; i < n; ++i) {
        g[i] = rec[i].g[k];
        batch[i] = rec[i].batch[k];
        const double R = g[i].trans ? g[i].RB : g[i].RA, C = g[i].trans ? g[i].RA : g[i].RB;
        flops += 2.0 * R * C * g[i].K * batch[i];
        gbytes += 4.0 * batch[i] * (R * g[i].K + R * C);
      }
      ProfScope ps(gemm_name(g[0]), flops, gbytes, st);
      launch_dst_gemm_group(g, batch, n, st);
    }
  };
  if (x6) {
    solves(true, 6, 6, 6, false, true, true);
  } else if (!mixed) {
    solves(true, NP_F32, NP_F32, NP_F32, false, true, true);
  } else {
    solves(true, 3, 6, 3, false, true, true);
    {
      ResidualArgs r[MAX_GROUP];
      for (int i = 0; i < n; ++i) r[i] = residual_args(*it[i].b, it[i].ws);
      ProfScope ps("residual", 0.0, 12.0 * bytes, st);
      launch_residual_group(r, n, st);
    }
    solves(false, 3, 1, 1, true, true, false);
  }
  LayerArgs la[MAX_GROUP];
  PlaneSumArgs sums[MAX_GROUP];
  int nm = 0;
  for (int i = 0; i < n; ++i) {
    la[i] = layer_args(*it[i].b, it[i].psi, it[i].ws, mixed);
    if (it[i].b->any_mass) {
      sums[nm] = mass_sums_args(*it[i].b, la[i], it[i].ws);
      mass_layer_args(*it[i].b, la[i], sums[nm], it[i].ws);
      ++nm;
    }
  }
  if (nm) {
    ProfScope ps("mass_constraint", 0.0, (mixed ? 8.0 : 4.0) * bytes, st);
    launch_plane_sums_group(sums, nm, st);
  }
  {
    ProfScope ps("to_layer", 0.0, (mixed ? 12.0 : 8.0) * bytes, st);
    launch_to_layer_group(la, n, st);
  }
}

// Parameters of a call: per member (device array) or one uniform set.
struct CallParams {
  const MemberParams* mp;
  MemberParams u;
  int biharm;
};

StageArgs stage_args(const Batch& p, const float* qs, const float* psi, const float* qn, float* out,
                     const CallParams& cp, float c1, float c2, bool apply_bc, void* ws) {
  StageArgs a{};
  a.qs = qs;
  a.psi = psi;
  a.qn = qn;
  a.out = out;
  a.Nx = p.Nx;
  a.Ny = p.Ny;
  a.P = p.P;
  a.plane = p.splane;
  a.dbeta_rows = p.dbeta_rows;
  a.wind_rows = p.wind_rows;
  a.t = p.t;
  if (p.any_mass) a.t.psi_wall = wsf(ws, p.off_PWALL);  // wall value of psi from the last inversion
  a.mp = cp.mp;
  a.u = cp.u;
  a.biharm = cp.biharm;
  a.idx2 = (float)(1.0 / (p.dx * p.dx));
  a.idy2 = (float)(1.0 / (p.dy * p.dy));
  a.iarakawa = (float)(1.0 / (12.0 * p.dx * p.dy));
  a.idx = (float)(1.0 / p.dx);
  a.idy = (float)(1.0 / p.dy);
  a.c1 = c1;
  a.c2 = c2;
  a.advection = p.advection;
  a.apply_bc = apply_bc ? 1 : 0;
  a.tile_rows = p.stage_rows;
  return a;
}

// algorithmic bytes of a stage: read qs and psi, read qn when combined, write out
double stage_bytes(const Batch& p, bool qn) { return 4.0 * p.P * (double)p.nx * p.ny * (3.0 + (qn ? 1.0 : 0.0)); }

void stage(const Batch& p, const float* qs, const float* psi, const float* qn, float* out, const CallParams& cp,
           float c1, float c2, bool apply_bc, void* ws, hipStream_t st) {
  const StageArgs a = stage_args(p, qs, psi, qn, out, cp, c1, c2, apply_bc, ws);
  ProfScope ps("qg_stage", 0.0, stage_bytes(p, qn != nullptr), st);
  launch_qg_stage(a, st);
}

// nsteps SSP-RK3 steps of the batch in place (Shu-Osher form, boundary
// condition before every RHS).
void rk3(const Batch& p, float* q, float* psi_last, int nsteps, double dt, const CallParams& cp, void* ws,
         hipStream_t st) {
  float* psi = psi_last ? psi_last : wsf(ws, p.off_PSI);
  float* q1 = wsf(ws, p.off_Q1);
  float* q2 = wsf(ws, p.off_Q2);
  const float h = (float)dt;
  for (int n = 0; n < nsteps; ++n) {
    invert(p, q, psi, ws, st);
    stage(p, q, psi, nullptr, q1, cp, 1.0f, h, true, ws, st);
    invert(p, q1, psi, ws, st);
    stage(p, q1, psi, q, q2, cp, 0.25f, 0.25f * h, true, ws, st);
    invert(p, q2, psi, ws, st);
    stage(p, q2, psi, q, q, cp, 2.0f / 3.0f, 2.0f / 3.0f * h, true, ws, st);
  }
}

// A batch of a merged SSP-RK3 group: its state and parameters.
struct RkItem {
  const Batch* b;
  float* q;
  CallParams cp;
  void* ws;
};

// rk3() for the batches of a group (same precision, advection, biharmonic
// flag and stage tile height): every kernel phase of every stage is one
// launch over the group.
void rk3_group(const std::vector<RkItem>& it, int nsteps, double dt, hipStream_t st) {
  const int n = (int)it.size();
  const float h = (float)dt;
  std::vector<GroupItem> gi(n);
  for (int i = 0; i < n; ++i) gi[i] = GroupItem{it[i].b, nullptr, wsf(it[i].ws, it[i].b->off_PSI), it[i].ws};
  auto q1 = [&](int i) { return wsf(it[i].ws, it[i].b->off_Q1); };
  auto q2 = [&](int i) { return wsf(it[i].ws, it[i].b->off_Q2); };
  // one stage: inversion of `in`, then out = qn + c1 (in - qn) + c2 rhs
  auto rk_stage = [&](auto in, auto qn, auto out, float c1, float c2) {
    for (int i = 0; i < n; ++i) gi[i].q = in(i);
    invert_group(gi, st);
    StageArgs a[MAX_GROUP];
    double bytes = 0.0;
    for (int i = 0; i < n; ++i) {
      a[i] = stage_args(*it[i].b, in(i), gi[i].psi, qn(i), out(i), it[i].cp, c1, c2, true, it[i].ws);
      bytes += stage_bytes(*it[i].b, qn(i) != nullptr);
    }
    ProfScope ps("qg_stage", 0.0, bytes, st);
    launch_qg_stage_group(a, n, st);
  };
  auto q0 = [&](int i) { return it[i].q; };
  auto none = [](int) -> float* { return nullptr; };
  for (int s = 0; s < nsteps; ++s) {
    rk_stage(q0, none, q1, 1.0f, h);
    rk_stage(q1, q0, q2, 0.25f, 0.25f * h);
    rk_stage(q2, q0, q0, 2.0f / 3.0f, 2.0f / 3.0f * h);
  }
}

// Tsit5 (FSAL): k0 = f(y) once, then per step six more evaluations
// k_i = f(y + dt sum_j a_ij k_j), i = 1..6, the last at the new state
// y + dt sum_j b_j k_j, which becomes k0 of the next step.  f = the RHS with
// the boundary condition applied first (SomaxModel.build_terms).
// tws: Y (stage state) then K0..K6, each P planes.
void tsit5(const Batch& p, float* q, int nsteps, double dt, const CallParams& cp, void* ws, void* tws,
           hipStream_t st) {
  const long n = (long)p.P * p.splane;
  float* psi = wsf(ws, p.off_PSI);
  float* Y = static_cast<float*>(tws);
  float* K[7];
  for (int i = 0; i < 7; ++i) K[i] = Y + (i + 1) * n;
  auto f = [&](const float* y, float* k) {
    invert(p, y, psi, ws, st);
    stage(p, y, psi, nullptr, k, cp, 0.0f, 1.0f, true, ws, st);
  };
  if (nsteps > 0) f(q, K[0]);
  for (int s = 0; s < nsteps; ++s) {
    for (int i = 1; i <= 6; ++i) {
      LinComb c{};
      c.y = q;
      c.nk = i;
      for (int j = 0; j < i; ++j) {
        c.k[j] = K[j];
        c.w[j] = dt * TSIT5_A[i][j];
      }
      c.n = n;
      c.out = i < 6 ? Y : q;  // row 6 = b: the new state, in place
      launch_lincomb(c, st);
      f(c.out, K[i]);
    }
    std::swap(K[0], K[6]);  // first same as last
  }
}

CallParams uniform_params(const somaxhip_qg_params* prm) {
  return CallParams{nullptr, to_member_params(*prm), prm->nu4 != 0.0 ? 1 : 0};
}

// An FP32 / BF16 plan and non-null buffers.
int check_plan(somaxhip_qg_plan p, const char* who) {
  if (p && p->f64) {
    set_error(std::string(who) + ": the plan computes in FP64 (SOMAXHIP_PREC_FP64); call the _f64 entry point");
    return SOMAXHIP_ERR_ARG;
  }
  return SOMAXHIP_OK;
}

int check_call(somaxhip_qg_plan p, const void* a, const void* b, void* ws, const char* who) {
  if (!p || !a || !b || !ws) {
    set_error(std::string(who) + ": null plan, buffer or workspace");
    return SOMAXHIP_ERR_ARG;
  }
  return check_plan(p, who);
}

// Sums the events of a profiled run into per-name totals.
int collect_profile(Profile& prof, hipStream_t st, somaxhip_kernel_stat* stats, int max_stats, int* nstats) {
  hipError_t e = hipStreamSynchronize(st);
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
  SOMAX_HIP_CHECK(e);
  return SOMAXHIP_OK;
}

}  // namespace

extern "C" {

int somaxhip_qg_invert(somaxhip_qg_plan p, const float* q, float* psi, int apply_bc, void* ws,
                       somaxhip_stream_t stream) {
  int rc = check_call(p, q, psi, ws, "somaxhip_qg_invert");
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  invert(p->b, q, psi, ws, st, apply_bc != 0);
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
  float* psi = psi_out ? psi_out : wsf(ws, p->b.off_PSI);
  invert(p->b, q, psi, ws, st, apply_bc != 0);
  stage(p->b, q, psi, nullptr, dq, uniform_params(prm), 0.0f, 1.0f, apply_bc != 0, ws, st);
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
  rk3(p->b, q, psi_out, nsteps, dt, uniform_params(prm), ws, st);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

}  // extern "C"

extern "C" {

int somaxhip_qg_step_noalloc(somaxhip_qg_plan p, float* q, float* psi_out, int nsteps, double dt,
                             const somaxhip_qg_params* prm, somaxhip_stream_t stream) {
  if (!p || !q || !prm || nsteps < 0) {
    set_error("somaxhip_qg_step_noalloc: null plan, state or params, or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  if (const int rc = check_plan(p, "somaxhip_qg_step_noalloc")) return rc;
  if (p->b.ws_bytes > (size_t)SOMAXHIP_NOALLOC_SCRATCH_BYTES) {
    set_error("somaxhip_qg_step_noalloc: the plan needs " + std::to_string(p->b.ws_bytes) +
              " bytes of workspace, more than the " + std::to_string((size_t)SOMAXHIP_NOALLOC_SCRATCH_BYTES) +
              " bytes of the library's scratch array; use somaxhip_qg_step with a caller workspace");
    return SOMAXHIP_ERR_ARG;
  }
  void* ws = nullptr;
  const int rc = scratch_address(&ws, "somaxhip_qg_step_noalloc");
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  rk3(p->b, q, psi_out, nsteps, dt, uniform_params(prm), ws, st);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_project(somaxhip_qg_plan p, double f0, const double* H, const double* M, const float* h,
                        const float* u, const float* v, int h_deviation, float* ho, float* uo, float* vo,
                        float* psi_out, void* ws, somaxhip_stream_t stream) {
  int rc = check_call(p, h, ho, ws, "somaxhip_qg_project");
  if (rc) return rc;
  if (!H || !M || !u || !v || !uo || !vo) {
    set_error("somaxhip_qg_project: null H, M or field");
    return SOMAXHIP_ERR_ARG;
  }
  const Batch& b = p->b;
  if (b.M != 1) {
    set_error("somaxhip_qg_project: needs a single-model plan");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  ProjArgs a{};
  a.h = h;
  a.u = u;
  a.v = v;
  a.q = wsf(ws, b.off_Q1);
  a.psi = psi_out ? psi_out : wsf(ws, b.off_PSI);
  a.ho = ho;
  a.uo = uo;
  a.vo = vo;
  a.Nx = b.Nx;
  a.Ny = b.Ny;
  a.nl = b.P;
  a.plane = b.splane;
  a.idx = (float)(1.0 / b.dx);
  a.idy = (float)(1.0 / b.dy);
  a.f0 = (float)f0;
  for (int k = 0; k < b.P; ++k) {
    a.H[k] = (float)H[k];
    for (int l = 0; l < b.P; ++l) a.M[k * MAX_LAYERS + l] = (float)M[k * b.P + l];
  }
  a.hdev = h_deviation ? 1 : 0;
  launch_project_pv(a, st);
  invert(b, a.q, const_cast<float*>(a.psi), ws, st, false);
  launch_project_geo(a, st);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

size_t somaxhip_qg_tsit5_workspace_bytes(somaxhip_qg_plan p) {
  if (!p) return 0;
  return p->f64 ? qg64_tsit5_workspace_bytes(p->f64) : (size_t)8 * p->b.P * p->b.splane * sizeof(float);
}

int somaxhip_qg_step_tsit5(somaxhip_qg_plan p, float* q, int nsteps, double dt, const somaxhip_qg_params* prm,
                           void* ws, void* tws, somaxhip_stream_t stream) {
  int rc = check_call(p, q, q, ws, "somaxhip_qg_step_tsit5");
  if (rc) return rc;
  if (!prm || !tws || nsteps < 0) {
    set_error("somaxhip_qg_step_tsit5: null params or Tsit5 workspace, or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  tsit5(p->b, q, nsteps, dt, uniform_params(prm), ws, tws, st);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_diagnose(somaxhip_qg_plan p, const float* q, float* psi, float* u, float* v, float* zeta,
                         double* ke_enstrophy, void* ws, somaxhip_stream_t stream) {
  int rc = check_call(p, q, psi, ws, "somaxhip_qg_diagnose");
  if (rc) return rc;
  if (!ke_enstrophy) {
    set_error("somaxhip_qg_diagnose: null ke_enstrophy");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const Batch& b = p->b;
  invert(b, q, psi, ws, st, false);
  QGDiagArgs a{};
  a.q = q;
  a.psi = psi;
  a.u = u;
  a.v = v;
  a.zeta = zeta;
  a.Nx = b.Nx;
  a.Ny = b.Ny;
  a.P = b.P;
  a.plane = b.splane;
  a.idx = (float)(1.0 / b.dx);
  a.idy = (float)(1.0 / b.dy);
  a.idx2 = (float)(1.0 / (b.dx * b.dx));
  a.idy2 = (float)(1.0 / (b.dy * b.dy));
  a.area = b.dx * b.dy;
  a.partial = reinterpret_cast<double*>(static_cast<char*>(ws) + b.off_DIAG);
  a.out = ke_enstrophy;
  launch_qg_diag(a, st);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_zero_ghost(float* f, int nl, int ny, int nx, somaxhip_stream_t stream) {
  if (!f || nl < 1 || nx < 1 || ny < 1) {
    set_error("somaxhip_zero_ghost: null field or bad shape");
    return SOMAXHIP_ERR_ARG;
  }
  launch_zero_ghost(f, nl, ny + 2, nx + 2, reinterpret_cast<hipStream_t>(stream));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_qg_step_profile(somaxhip_qg_plan p, float* q, int nsteps, double dt, const somaxhip_qg_params* prm,
                             void* ws, somaxhip_stream_t stream, somaxhip_kernel_stat* stats, int max_stats,
                             int* nstats) {
  if (!stats || !nstats || max_stats < 1) {
    set_error("somaxhip_qg_step_profile: null stats");
    return SOMAXHIP_ERR_ARG;
  }
  Profile prof;
  g_prof = &prof;
  const int rc = somaxhip_qg_step(p, q, nullptr, nsteps, dt, prm, ws, stream);
  g_prof = nullptr;
  const int rc2 = collect_profile(prof, reinterpret_cast<hipStream_t>(stream), stats, max_stats, nstats);
  return rc ? rc : rc2;
}

// ---------------------------------------------------------------------------
// Ensembles
// ---------------------------------------------------------------------------

int somaxhip_ens_create(const somaxhip_qg_desc* members, const somaxhip_qg_params* params, int n, int flags,
                        somaxhip_ens* out) {
  if (!members || !params || !out || n < 1) {
    set_error("somaxhip_ens_create: null argument or no members");
    return SOMAXHIP_ERR_ARG;
  }
  *out = nullptr;
  if (flags & ~(SOMAXHIP_ENS_CONCURRENT | SOMAXHIP_ENS_GRAPH | SOMAXHIP_ENS_MERGED)) {
    set_error("somaxhip_ens_create: unknown flags");
    return SOMAXHIP_ERR_ARG;
  }
  for (int i = 0; i < n; ++i) {
    const std::string bad = check_desc(&members[i], false);
    if (!bad.empty()) {
      set_error("somaxhip_ens_create: member " + std::to_string(i) + ": " + bad);
      return SOMAXHIP_ERR_ARG;
    }
  }
  std::unique_ptr<somaxhip_ens_s> E(new (std::nothrow) somaxhip_ens_s());
  if (!E) {
    set_error("somaxhip_ens_create: out of host memory");
    return SOMAXHIP_ERR_ALLOC;
  }
  E->flags = flags;
  // group members into classes of one grid; classes ordered by plane size,
  // largest first, so the longest launches are queued first
  std::vector<int> rep;  // class -> representative member
  E->mclass.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    int c = 0;
    while (c < (int)rep.size() && !same_class(&members[rep[c]], &members[i])) ++c;
    if (c == (int)rep.size()) rep.push_back(i);
    E->mclass[i] = c;
  }
  const int C = (int)rep.size();
  std::vector<int> order(C);
  for (int c = 0; c < C; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const long sa = (long)(members[rep[a]].nx + 2) * (members[rep[a]].ny + 2);
    const long sb = (long)(members[rep[b]].nx + 2) * (members[rep[b]].ny + 2);
    return sa > sb;
  });
  std::vector<int> rank(C);
  for (int k = 0; k < C; ++k) rank[order[k]] = k;
  for (int i = 0; i < n; ++i) E->mclass[i] = rank[E->mclass[i]];
  E->cmembers.assign(C, {});
  for (int i = 0; i < n; ++i) E->cmembers[E->mclass[i]].push_back(i);

  E->moff.assign(n, 0);
  size_t state = 0, ws = 0;
  for (int c = 0; c < C; ++c) {
    std::vector<const somaxhip_qg_desc*> ds;
    for (int i : E->cmembers[c]) ds.push_back(&members[i]);
    std::unique_ptr<Batch> b(new (std::nothrow) Batch());
    if (!b) {
      set_error("somaxhip_ens_create: out of host memory");
      return SOMAXHIP_ERR_ALLOC;
    }
    hipError_t e = batch_build(*b, ds);
    MemberParams* dp = nullptr;
    std::vector<MemberParams> hp;
    int biharm = 0;
    for (int i : E->cmembers[c]) {
      hp.push_back(to_member_params(params[i]));
      biharm |= params[i].nu4 != 0.0;
    }
    if (e == hipSuccess) e = b->mem.upload(&dp, hp);
    if (e != hipSuccess) {
      set_error(std::string("somaxhip_ens_create: ") + hipGetErrorString(e));
      return SOMAXHIP_ERR_ALLOC;
    }
    E->cstate.push_back(state);
    E->cws.push_back(ws);
    for (int k = 0; k < b->M; ++k) E->moff[E->cmembers[c][k]] = (long)(state + (size_t)b->first[k] * b->splane);
    state += (size_t)b->P * b->splane;
    state = (size_t)round_up((long)state, 64);  // 256-byte aligned class blocks
    ws += (size_t)round_up((long)b->ws_bytes, 256);
    E->cparams.push_back(dp);
    E->hparams.push_back(hp);
    E->cbiharm.push_back(biharm);
    E->cls.push_back(std::move(b));
  }
  E->state_floats = state;
  E->ws_bytes = ws;
  // streams (default priority) and events for the class branches
  SOMAX_HIP_CHECK(hipStreamCreateWithFlags(&E->main, hipStreamNonBlocking));
  SOMAX_HIP_CHECK(hipEventCreateWithFlags(&E->ev_in, hipEventDisableTiming));
  SOMAX_HIP_CHECK(hipEventCreateWithFlags(&E->ev_out, hipEventDisableTiming));
  SOMAX_HIP_CHECK(hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming));
  for (int c = 0; c < C; ++c) {
    hipStream_t s;
    hipEvent_t ev;
    SOMAX_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    E->side.push_back(s);
    SOMAX_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    E->ev_join.push_back(ev);
  }
  *out = E.release();
  return SOMAXHIP_OK;
}

int somaxhip_ens_destroy(somaxhip_ens e) {
  delete e;
  return SOMAXHIP_OK;
}

size_t somaxhip_ens_state_floats(somaxhip_ens e) { return e ? e->state_floats : 0; }

size_t somaxhip_ens_workspace_bytes(somaxhip_ens e) { return e ? e->ws_bytes : 0; }

long somaxhip_ens_member_offset(somaxhip_ens e, int member) {
  if (!e || member < 0 || member >= (int)e->moff.size()) {
    set_error("somaxhip_ens_member_offset: bad ensemble or member index");
    return -1;
  }
  return e->moff[member];
}

int somaxhip_ens_num_classes(somaxhip_ens e) { return e ? (int)e->cls.size() : 0; }

int somaxhip_ens_set_params(somaxhip_ens e, const somaxhip_qg_params* params, somaxhip_stream_t stream) {
  if (!e || !params) {
    set_error("somaxhip_ens_set_params: null argument");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (size_t c = 0; c < e->cls.size(); ++c) {
    int biharm = 0;
    for (size_t k = 0; k < e->cmembers[c].size(); ++k) {
      const somaxhip_qg_params& p = params[e->cmembers[c][k]];
      e->hparams[c][k] = to_member_params(p);
      biharm |= p.nu4 != 0.0;
    }
    if (biharm != e->cbiharm[c]) e->stale = true;  // the captured step picked the other stage kernel
    e->cbiharm[c] = biharm;
    SOMAX_HIP_CHECK(hipMemcpyAsync(e->cparams[c], e->hparams[c].data(), e->hparams[c].size() * sizeof(MemberParams),
                                   hipMemcpyHostToDevice, st));
  }
  // the pageable source is reused by the next call: wait for the copies
  SOMAX_HIP_CHECK(hipStreamSynchronize(st));
  return SOMAXHIP_OK;
}

int somaxhip_ens_class_info(somaxhip_ens e, int cls, int* nx, int* ny, int* planes, int* members, int* fold,
                            int* stage_rows) {
  if (!e || cls < 0 || cls >= (int)e->cls.size()) {
    set_error("somaxhip_ens_class_info: null ensemble or class out of range");
    return SOMAXHIP_ERR_ARG;
  }
  const Batch& b = *e->cls[cls];
  if (nx) *nx = b.nx;
  if (ny) *ny = b.ny;
  if (planes) *planes = b.P;
  if (members) *members = b.M;
  if (fold) *fold = b.fold;
  if (stage_rows) *stage_rows = b.stage_rows;
  return SOMAXHIP_OK;
}

int somaxhip_ens_set_stage_rows(somaxhip_ens e, int cls, int rows) {
  if (!e || cls < 0 || cls >= (int)e->cls.size() || (rows != 8 && rows != 16)) {
    set_error("somaxhip_ens_set_stage_rows: null ensemble, class out of range or rows not 8 or 16");
    return SOMAXHIP_ERR_ARG;
  }
  if (e->stepped) {
    set_error("somaxhip_ens_set_stage_rows: the tile height is fixed once the ensemble has stepped");
    return SOMAXHIP_ERR_ARG;
  }
  e->cls[cls]->stage_rows = rows;
  return SOMAXHIP_OK;
}

}  // extern "C"

namespace {

// Launch groups of a merged ensemble: classes in order, each joining the
// first group with the same precision, advection, biharmonic flag and (for
// the Arakawa stage) tile height that still has room.
std::vector<std::vector<int>> ens_groups(somaxhip_ens e) {
  std::vector<std::vector<int>> groups;
  auto key_eq = [&](int a, int b) {
    const Batch& x = *e->cls[a];
    const Batch& y = *e->cls[b];
    return x.precision == y.precision && x.advection == y.advection && e->cbiharm[a] == e->cbiharm[b] &&
           (x.advection == 1 || x.stage_rows == y.stage_rows);
  };
  for (int c = 0; c < (int)e->cls.size(); ++c) {
    size_t g = 0;
    while (g < groups.size() && !(groups[g].size() < (size_t)MAX_GROUP && key_eq(groups[g][0], c))) ++g;
    if (g == groups.size()) groups.emplace_back();
    groups[g].push_back(c);
  }
  return groups;
}

// One ensemble step queued behind `st`.  Per class (default): each class
// advances with its own launches, on its own stream when concurrent (fork /
// join through events, which stream capture turns into parallel graph
// branches), else all on `st`.  Merged: each launch group advances with one
// launch per kernel phase, the groups on their own streams when concurrent.
int ens_queue(somaxhip_ens e, float* q, int nsteps, double dt, void* ws, hipStream_t st) {
  const int C = (int)e->cls.size();
  // every per-class table is indexed by the class below
  const size_t Cs = (size_t)C;
  if (e->cparams.size() != Cs || e->cbiharm.size() != Cs || e->cstate.size() != Cs || e->cws.size() != Cs ||
      e->side.size() != Cs || e->ev_join.size() != Cs) {
    set_error("somaxhip_ens_step: inconsistent ensemble tables");
    return SOMAXHIP_ERR_ARG;
  }
  const bool merged = (e->flags & SOMAXHIP_ENS_MERGED) != 0;
  const std::vector<std::vector<int>> groups = merged ? ens_groups(e) : std::vector<std::vector<int>>{};
  const int lanes = merged ? (int)groups.size() : C;  // independent chains (<= C)
  const bool conc = (e->flags & SOMAXHIP_ENS_CONCURRENT) && lanes > 1;
  if (conc) {
    SOMAX_HIP_CHECK(hipEventRecord(e->ev_fork, st));
    for (int c = 0; c < lanes; ++c) SOMAX_HIP_CHECK(hipStreamWaitEvent(e->side[c], e->ev_fork, 0));
  }
  for (int c = 0; c < lanes; ++c) {
    hipStream_t s = conc ? e->side[c] : st;
    if (merged) {
      std::vector<RkItem> it;
      for (int k : groups[c]) {
        const CallParams cp{e->cparams[k], MemberParams{0, 0, 0, 0}, e->cbiharm[k]};
        it.push_back(RkItem{e->cls[k].get(), q + e->cstate[k], cp, static_cast<char*>(ws) + e->cws[k]});
      }
      rk3_group(it, nsteps, dt, s);
    } else {
      const Batch& b = *e->cls[c];
      const CallParams cp{e->cparams[c], MemberParams{0, 0, 0, 0}, e->cbiharm[c]};
      rk3(b, q + e->cstate[c], nullptr, nsteps, dt, cp, static_cast<char*>(ws) + e->cws[c], s);
    }
  }
  if (conc) {
    for (int c = 0; c < lanes; ++c) {
      SOMAX_HIP_CHECK(hipEventRecord(e->ev_join[c], e->side[c]));
      SOMAX_HIP_CHECK(hipStreamWaitEvent(st, e->ev_join[c], 0));
    }
  }
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

}  // namespace

extern "C" {

int somaxhip_ens_step(somaxhip_ens e, float* q, int nsteps, double dt, void* ws, somaxhip_stream_t stream) {
  if (!e || !q || !ws || nsteps < 0) {
    set_error("somaxhip_ens_step: null ensemble, state or workspace, or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (nsteps == 0) return SOMAXHIP_OK;
  e->stepped = true;
  if (!(e->flags & SOMAXHIP_ENS_GRAPH)) return ens_queue(e, q, nsteps, dt, ws, st);

  // capture one step on the private stream and replay it there, ordered
  // after the caller's stream and before its later work by two events
  if (!e->exec || e->stale || e->g_q != q || e->g_ws != ws || e->g_dt != dt) {
    if (e->exec) {
      SOMAX_HIP_CHECK(hipStreamSynchronize(e->main));  // replays of the old graph may be queued
      (void)hipGraphExecDestroy(e->exec);
      e->exec = nullptr;
    }
    e->stale = false;
    hipGraph_t graph = nullptr;
    SOMAX_HIP_CHECK(hipStreamBeginCapture(e->main, hipStreamCaptureModeThreadLocal));
    const int rc = ens_queue(e, q, 1, dt, ws, e->main);
    const hipError_t ec = hipStreamEndCapture(e->main, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    SOMAX_HIP_CHECK(ec);
    const hipError_t ei = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SOMAX_HIP_CHECK(ei);
    e->g_q = q;
    e->g_ws = ws;
    e->g_dt = dt;
  }
  SOMAX_HIP_CHECK(hipEventRecord(e->ev_in, st));
  SOMAX_HIP_CHECK(hipStreamWaitEvent(e->main, e->ev_in, 0));
  for (int n = 0; n < nsteps; ++n) SOMAX_HIP_CHECK(hipGraphLaunch(e->exec, e->main));
  SOMAX_HIP_CHECK(hipEventRecord(e->ev_out, e->main));
  SOMAX_HIP_CHECK(hipStreamWaitEvent(st, e->ev_out, 0));
  return SOMAXHIP_OK;
}

int somaxhip_ens_step_profile(somaxhip_ens e, float* q, int nsteps, double dt, void* ws, somaxhip_stream_t stream,
                              somaxhip_kernel_stat* stats, int max_stats, int* nstats) {
  if (!e || !q || !ws || nsteps < 0 || !stats || !nstats || max_stats < 1) {
    set_error("somaxhip_ens_step_profile: null argument or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  // serial and eager, so every launch is timed on the caller's stream (merged
  // launch groups stay merged)
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  e->stepped = true;
  const int saved = e->flags;
  e->flags = saved & SOMAXHIP_ENS_MERGED;
  Profile prof;
  g_prof = &prof;
  const int rc = ens_queue(e, q, nsteps, dt, ws, st);
  g_prof = nullptr;
  e->flags = saved;
  const int rc2 = collect_profile(prof, st, stats, max_stats, nstats);
  return rc ? rc : rc2;
}

}  // extern "C"
