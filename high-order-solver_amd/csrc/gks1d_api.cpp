// gks1d_api.cpp — the extern "C" layer of include/gks1d.h over the persistent
// 1D kernel.  One context = one tube = one HIP stream; not re-entrant.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gks1d.h"
#include "gks1d_launch.hpp"

namespace {

thread_local std::string g_err1;

int fail(int code, const char* what) {
  g_err1 = what;
  return code;
}
int hip_fail(hipError_t e, const char* where) {
  g_err1 = std::string(where) + ": " + hipGetErrorString(e);
  return GKS_ERR_HIP;
}
#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

gksd::Gas gas_of(const gks1d_params& p) {
  gksd::Gas g;
  g.K = p.K;
  g.gamma = p.gamma;
  g.mu = p.mu;
  g.nu = p.nu;
  g.pr = p.pr;
  g.c1 = p.c1_euler;
  g.c2 = p.c2_euler;
  g.tau_type = p.tau_type;
  g.prandtl_fix = 0;
  g.inv_gm1 = 1.0 / (p.gamma - 1.0);
  return g;
}

int check_params(const gks1d_params* p) {
  if (!p) return fail(GKS_ERR_ARG, "null params");
  if (p->ghost != 5) return fail(GKS_ERR_UNSUPPORTED, "ghost must be 5");
  if (p->nodex < 5) return fail(GKS_ERR_ARG, "nodex must be >= 5");
  if (p->recon_variable != GKS_VAR_CONSERVATIVE && p->recon_variable != GKS_VAR_CHARACTERISTIC)
    return fail(GKS_ERR_ARG, "recon_variable");
  if (p->use_df != 0 && p->use_df != 1) return fail(GKS_ERR_ARG, "use_df");
  if (p->df_from_start != 0 && p->df_from_start != 1) return fail(GKS_ERR_ARG, "df_from_start");
  if (p->tau_type != GKS_TAU_EULER && p->tau_type != GKS_TAU_NS) return fail(GKS_ERR_ARG, "tau_type");
  if (!(p->right > p->left)) return fail(GKS_ERR_ARG, "empty domain");
  if (!(p->gamma > 1.0) || p->K < 0) return fail(GKS_ERR_ARG, "gamma / K");
  if (!(p->cfl > 0.0)) return fail(GKS_ERR_ARG, "cfl");
  return GKS_OK;
}

// geometry and gas of a tube (1D SetUniformMesh, basic_function.cpp:1262-1309);
// the buffer pointers are left null
void dev_of(const gks1d_params& p, Dev1& d) {
  std::memset(&d, 0, sizeof(d));
  d.N = p.nodex;
  d.G = p.ghost;
  d.nx = d.N + 2 * d.G;
  d.recon_var = p.recon_variable;
  d.use_df = p.use_df;
  d.df_from_start = p.df_from_start;
  d.dx = (p.right - p.left) / p.nodex;
  d.idx = 1.0 / d.dx;
  d.cfl = p.cfl;
  d.cfl_dx = p.cfl * d.dx;                        // CFL * dx (Dtx, time_advance.cpp:93-111)
  d.visc_dt = 0.25 * p.cfl * d.dx * d.dx / p.mu;  // the viscous limit of Dtx
  d.gas = gas_of(p);
}

int check_run_args(const Dev1& d, const int32_t bc[2], int32_t max_steps) {
  for (int s = 0; s < 2; s++)
    if (bc[s] < 0 || bc[s] > 2) return fail(GKS_ERR_ARG, "bc kind");
  if (max_steps < 0 || max_steps > 10000000) return fail(GKS_ERR_ARG, "max_steps must be in [0, 1e7]");
  if ((bc[0] != GKS_BC_FREE || bc[1] != GKS_BC_FREE) && d.N < d.G)
    return fail(GKS_ERR_ARG, "periodic / reflective boundaries need nodex >= ghost");
  return GKS_OK;
}

// The driver's time loop in launches of the persistent kernel (up to
// kStepsPerLaunch1d steps each); d_dt is the device dt record of one launch.
int run_loop(const Dev1& d, const Launch1dTable* lt, hipStream_t s, double* d_dt, const int32_t bc[2],
             double tstop, int32_t max_steps, double* t_inout, int32_t* step_inout, double* dt_hist) {
  double t = *t_inout;
  int32_t step = *step_inout;
  int32_t n = 0;
  double st[4];
  while (t < tstop && n < max_steps) {
    const int chunk = max_steps - n < kStepsPerLaunch1d ? max_steps - n : kStepsPerLaunch1d;
    HIPCHK(lt->run(d, bc[0], bc[1], tstop, t, step, chunk, d_dt, s));
    HIPCHK(hipMemcpyAsync(st, d.st, sizeof(st), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int ran = (int)st[2];
    if (dt_hist && ran > 0) {
      HIPCHK(hipMemcpyAsync(dt_hist + n, d_dt, (size_t)ran * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    t = st[0];
    step = (int32_t)st[1];
    n += ran;
    if (ran < chunk) break;  // the loop condition ended the launch
  }
  *t_inout = t;
  *step_inout = step;
  return GKS_OK;
}

}  // namespace

// ---- the allocation-free run (gks1d_run_noalloc) -------------------------
// Every device buffer of the tube other than the caller's state (U_old, the
// DF sensor when the caller passes none, face states, face fluxes, the loop
// status and the dt record of one launch) lives in one file-scope array,
// sized for tubes of up to kNoAllocMaxNodex interior cells.  Each GPU holds
// its own copy for the life of the process; every call on a GPU shares it.
namespace {
constexpr size_t kNoAllocMaxNodex = 16384;
constexpr size_t kRnd = 32;  // doubles: each part starts on a 256-byte boundary
constexpr size_t rnd(size_t n) { return (n + kRnd - 1) / kRnd * kRnd; }
constexpr size_t kNxMax = kNoAllocMaxNodex + 2 * 5, kNfMax = kNxMax + 1;
constexpr size_t kOffUold = 0;
constexpr size_t kOffAlpha = kOffUold + rnd(3 * kNxMax);
constexpr size_t kOffFace = kOffAlpha + rnd(kNxMax);
constexpr size_t kOffFl = kOffFace + rnd(12 * kNfMax);
constexpr size_t kOffSt = kOffFl + rnd(12 * kNfMax);
constexpr size_t kOffDt = kOffSt + rnd(4);
constexpr size_t kScratchDoubles = kOffDt + rnd(kStepsPerLaunch1d);  // 3.71 MB
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current GPU's copy of `scratch`, looked up once
// per device with that device current.  Host code uses only this address:
// the array's own name is a host-side stand-in, not a device pointer.
int scratch_of_current_device(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(GKS_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(GKS_ERR_HIP, "hipGetSymbolAddress returned null");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return GKS_OK;
}

// zero fill (a kernel rather than hipMemsetAsync, so a caller's stream
// capture records no memset node)
__global__ void k_zero1d(double* p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}
}  // namespace

struct gks1d_ctx {
  gks1d_params p;
  Dev1 d;
  const Launch1dTable* lt = nullptr;
  hipStream_t s = nullptr;
  double* d_dt = nullptr;  // per-launch dt record (kStepsPerLaunch1d)
  std::vector<void*> allocs;
};

static int dev_alloc(gks1d_ctx* c, double** out, size_t n) {
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, n * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  c->allocs.push_back(ptr);
  e = hipMemset(ptr, 0, n * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "hipMemset");
  *out = static_cast<double*>(ptr);
  return GKS_OK;
}

extern "C" {

const char* gks1d_last_error(void) { return g_err1.c_str(); }

int gks1d_destroy(gks1d_ctx* c) {
  if (!c) return GKS_OK;
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (void* a : c->allocs) (void)hipFree(a);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return GKS_OK;
}

int gks1d_create(const gks1d_params* p, gks1d_ctx** out) {
  if (!out) return fail(GKS_ERR_ARG, "null out");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  gks1d_ctx* c = new gks1d_ctx();
  c->p = *p;
  Dev1& d = c->d;
  dev_of(*p, d);
  c->lt = p->fast_math ? &fast::table1d : &ieee::table1d;
  hipError_t e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    gks1d_destroy(c);
    return hip_fail(e, "hipStreamCreate");
  }
  const size_t nx = d.nx, nf = d.nx + 1;
  if ((rc = dev_alloc(c, &d.U, 3 * nx)) || (rc = dev_alloc(c, &d.Uold, 3 * nx)) ||
      (rc = dev_alloc(c, &d.alpha, nx)) || (rc = dev_alloc(c, &d.face, 12 * nf)) ||
      (rc = dev_alloc(c, &d.fl, 12 * nf)) || (rc = dev_alloc(c, &d.st, 4)) ||
      (rc = dev_alloc(c, &c->d_dt, kStepsPerLaunch1d))) {
    gks1d_destroy(c);
    return rc;
  }
  *out = c;
  return GKS_OK;
}

static void gather(std::vector<double>& soa, const double* base, long long n, int nvar, int64_t cs, int64_t vs) {
  const char* b = reinterpret_cast<const char*>(base);
  for (long long k = 0; k < n; k++)
    for (int v = 0; v < nvar; v++) soa[v * n + k] = *reinterpret_cast<const double*>(b + k * cs + v * vs);
}
static void scatter(const std::vector<double>& soa, double* base, long long n, int nvar, int64_t cs, int64_t vs) {
  char* b = reinterpret_cast<char*>(base);
  for (long long k = 0; k < n; k++)
    for (int v = 0; v < nvar; v++) *reinterpret_cast<double*>(b + k * cs + v * vs) = soa[v * n + k];
}

int gks1d_set_convar(gks1d_ctx* c, const double* base, int64_t cs, int64_t vs) {
  if (!c || !base) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(3 * (size_t)c->d.nx);
  gather(h, base, c->d.nx, 3, cs, vs);
  HIPCHK(hipMemcpyAsync(c->d.U, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return GKS_OK;
}
int gks1d_get_convar(gks1d_ctx* c, double* base, int64_t cs, int64_t vs) {
  if (!c || !base) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(3 * (size_t)c->d.nx);
  HIPCHK(hipMemcpyAsync(h.data(), c->d.U, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  scatter(h, base, c->d.nx, 3, cs, vs);
  return GKS_OK;
}
int gks1d_get_alpha(gks1d_ctx* c, double* alpha, int64_t cs) {
  if (!c || !alpha) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(c->d.nx);
  HIPCHK(hipMemcpyAsync(h.data(), c->d.alpha, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  scatter(h, alpha, c->d.nx, 1, cs, 0);
  return GKS_OK;
}
int gks1d_get_face_fluxes(gks1d_ctx* c, int32_t stage, double* out) {
  if (!c || !out) return fail(GKS_ERR_ARG, "null");
  if (stage < 0 || stage > 1) return fail(GKS_ERR_ARG, "stage");
  const size_t nf = c->d.nx + 1;
  std::vector<double> h(6 * nf);
  HIPCHK(hipMemcpyAsync(h.data(), c->d.fl + stage * 6 * nf, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (size_t i = 0; i < nf; i++)
    for (int v = 0; v < 6; v++) out[i * 6 + v] = h[v * nf + i];
  return GKS_OK;
}

int gks1d_run(gks1d_ctx* c, const int32_t bc[2], double tstop, int32_t max_steps, double* t_inout,
              int32_t* step_inout, double* dt_hist) {
  if (!c || !bc || !t_inout || !step_inout) return fail(GKS_ERR_ARG, "null");
  int rc = check_run_args(c->d, bc, max_steps);
  if (rc) return rc;
  return run_loop(c->d, c->lt, c->s, c->d_dt, bc, tstop, max_steps, t_inout, step_inout, dt_hist);
}

int gks1d_boundary(gks1d_ctx* c, int32_t side, int32_t kind) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (side < 0 || side > 1) return fail(GKS_ERR_ARG, "side");
  if (kind < 0 || kind > 2) return fail(GKS_ERR_ARG, "bc kind");
  if (kind != GKS_BC_FREE && c->d.N < c->d.G) return fail(GKS_ERR_ARG, "periodic / reflective boundaries need nodex >= ghost");
  HIPCHK(c->lt->phase(c->d, PH1_BC, side, kind, 0, 0.0, 0.0, 0.0, c->s));
  return GKS_OK;
}
int gks1d_copy_old(gks1d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(c->lt->phase(c->d, PH1_COPY_OLD, 0, 0, 0, 0.0, 0.0, 0.0, c->s));
  return GKS_OK;
}
int gks1d_cfl(gks1d_ctx* c, double t, double tstop, double* dt) {
  if (!c || !dt) return fail(GKS_ERR_ARG, "null");
  double st[4];
  HIPCHK(c->lt->phase(c->d, PH1_CFL, 0, 0, 0, 0.0, t, tstop, c->s));
  HIPCHK(hipMemcpyAsync(st, c->d.st, sizeof(st), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  *dt = st[3];
  return GKS_OK;
}
int gks1d_reconstruct(gks1d_ctx* c, int32_t with_df) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(c->lt->phase(c->d, PH1_RECON, with_df ? 1 : 0, 0, 0, 0.0, 0.0, 0.0, c->s));
  return GKS_OK;
}
int gks1d_flux(gks1d_ctx* c, int32_t stage, double dt) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (stage < 0 || stage > 1) return fail(GKS_ERR_ARG, "stage");
  HIPCHK(c->lt->phase(c->d, PH1_FLUX, 0, 0, stage, dt, 0.0, 0.0, c->s));
  return GKS_OK;
}
int gks1d_update(gks1d_ctx* c, int32_t stage) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (stage < 0 || stage > 1) return fail(GKS_ERR_ARG, "stage");
  HIPCHK(c->lt->phase(c->d, PH1_UPDATE, 0, 0, stage, 0.0, 0.0, 0.0, c->s));
  return GKS_OK;
}
int gks1d_update_alpha(gks1d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(c->lt->phase(c->d, PH1_ALPHA, 0, 0, 0, 0.0, 0.0, 0.0, c->s));
  return GKS_OK;
}

int gks1d_flux_points(const gks1d_params* p, int64_t n, const double* left, const double* right, double dt,
                      double* center, double* f, double* derf) {
  if (!p || !left || !right || !f || !derf) return fail(GKS_ERR_ARG, "null");
  if (n < 0) return fail(GKS_ERR_ARG, "n < 0");
  if (n == 0) return GKS_OK;
  const Launch1dTable* lt = p->fast_math ? &fast::table1d : &ieee::table1d;
  double *dl = nullptr, *dr = nullptr, *dc = nullptr, *dfo = nullptr, *dd = nullptr;
  hipError_t e = hipMalloc(&dl, n * 6 * 8);
  if (e == hipSuccess) e = hipMalloc(&dr, n * 6 * 8);
  if (e == hipSuccess) e = hipMalloc(&dc, n * 6 * 8);
  if (e == hipSuccess) e = hipMalloc(&dfo, n * 3 * 8);
  if (e == hipSuccess) e = hipMalloc(&dd, n * 3 * 8);
  if (e == hipSuccess) e = hipMemcpy(dl, left, n * 6 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dr, right, n * 6 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = lt->flux_points(gas_of(*p), n, dl, dr, dt, dc, dfo, dd, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && center) e = hipMemcpy(center, dc, n * 6 * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(f, dfo, n * 3 * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(derf, dd, n * 3 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(dl);
  (void)hipFree(dr);
  (void)hipFree(dc);
  (void)hipFree(dfo);
  (void)hipFree(dd);
  if (e != hipSuccess) return hip_fail(e, "gks1d_flux_points");
  return GKS_OK;
}

int gks1d_run_noalloc(const gks1d_params* p, double* U, double* alpha, const int32_t bc[2], double tstop,
                      int32_t max_steps, double* t_inout, int32_t* step_inout, double* dt_hist, void* stream) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!U || !bc || !t_inout || !step_inout) return fail(GKS_ERR_ARG, "null");
  if ((size_t)p->nodex > kNoAllocMaxNodex)
    return fail(GKS_ERR_ARG, "gks1d_run_noalloc: nodex exceeds the scratch limit of 16384 interior cells");
  Dev1 d;
  dev_of(*p, d);
  if ((rc = check_run_args(d, bc, max_steps))) return rc;
  double* scr = nullptr;
  if ((rc = scratch_of_current_device(&scr))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  d.U = U;
  d.Uold = scr + kOffUold;
  d.face = scr + kOffFace;
  d.fl = scr + kOffFl;
  d.st = scr + kOffSt;
  if (alpha) {
    d.alpha = alpha;
  } else {
    // a fresh DF sensor, as gks1d_create leaves it (ghost cells stay 0)
    d.alpha = scr + kOffAlpha;
    hipLaunchKernelGGL(k_zero1d, dim3((d.nx + 255) / 256), dim3(256), 0, s, d.alpha, d.nx);
    HIPCHK(hipGetLastError());
  }
  return run_loop(d, p->fast_math ? &fast::table1d : &ieee::table1d, s, scr + kOffDt, bc, tstop, max_steps,
                  t_inout, step_inout, dt_hist);
}

}  // extern "C"
