// gks3d_api.cpp — the extern "C" layer of include/gks3d.h over the 3D
// kernels.  One context = one cube = one HIP stream; not re-entrant.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gks3d.h"
#include "../../include/gks_learn.h"
#include "gks3d_launch.hpp"
#include "mlp_pack.hpp"

namespace {

thread_local std::string g_err3;

int fail(int code, const char* what) {
  g_err3 = what;
  return code;
}
int hip_fail(hipError_t e, const char* where) {
  g_err3 = std::string(where) + ": " + hipGetErrorString(e);
  return GKS_ERR_HIP;
}
#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)
#define RC(call)             \
  do {                       \
    int rc_ = (call);        \
    if (rc_) return rc_;     \
  } while (0)

gksd3::Gas3 gas_of(const gks3d_params& p) {
  gksd3::Gas3 g;
  g.K = p.K;
  g.gamma = p.gamma;
  g.mu = p.mu;
  g.nu = p.nu;
  g.c1 = p.c1_euler;
  g.c2 = p.c2_euler;
  g.tau_type = p.tau_type;
  return g;
}

int check_params(const gks3d_params* p) {
  if (!p) return fail(GKS_ERR_ARG, "null params");
  if (p->ghost < 3 || p->ghost > 8) return fail(GKS_ERR_UNSUPPORTED, "ghost must be in [3, 8]");
  if (p->nodex < 1 || p->nodey < 1 || p->nodez < 1) return fail(GKS_ERR_ARG, "nodex, nodey, nodez must be >= 1");
  const double n = (double)(p->nodex + 2 * p->ghost + 1) * (p->nodey + 2 * p->ghost + 1) * (p->nodez + 2 * p->ghost + 1);
  if (n * 60.0 > 9.0e18) return fail(GKS_ERR_ARG, "grid too large");
  if (p->recon_variable != GKS_VAR_CONSERVATIVE && p->recon_variable != GKS_VAR_CHARACTERISTIC)
    return fail(GKS_ERR_ARG, "recon_variable");
  if (p->tau_type != GKS_TAU_EULER && p->tau_type != GKS_TAU_NS) return fail(GKS_ERR_ARG, "tau_type");
  if (!(p->right > p->left) || !(p->up > p->down) || !(p->front > p->back)) return fail(GKS_ERR_ARG, "empty domain");
  if (!(p->gamma > 1.0) || p->K < 0) return fail(GKS_ERR_ARG, "gamma / K");
  if (!(p->cfl > 0.0)) return fail(GKS_ERR_ARG, "cfl");
  return GKS_OK;
}

}  // namespace

struct gks3d_ctx {
  gks3d_params p;
  Dev3 d;
  const Launch3dTable* lt = nullptr;
  hipStream_t s = nullptr;
  double* h_stat = nullptr;  // pinned [4]
  bool have_speed = false;   // CFL inputs known (set_cfl_primvar or a step)
  bool stat_pending = false; // a primitive sweep's reduction waits in d.stat
  double tmax = 0.0, rhomin = 0.0;
  std::vector<void*> allocs;
  // learned path (gks3d_set_classifier)
  float* d_mlp = nullptr;
  unsigned short* d_mlpw = nullptr;
  unsigned char* d_lin = nullptr;         // lin3 [3][nc], then lin_line [2][nf]
  unsigned long long* d_lstat = nullptr;  // [3][kStat3Slots][2]
  // per-phase HIP-event timing (gks3d_timing)
  unsigned timing = 0;  // bit mask of the timed phases
  struct Rec {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  double t_ms[GKS3_NPHASE] = {};
  long long t_n[GKS3_NPHASE] = {};
};

static hipEvent_t pool_event(gks3d_ctx* c) {
  if (c->pool_used == c->pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->pool.push_back(e);
  }
  return c->pool[c->pool_used++];
}
// launches one kernel phase; with timing on, brackets it with events on the
// context's stream
template <class F>
static int timed(gks3d_ctx* c, int phase, F&& launch) {
  hipEvent_t a = nullptr, b = nullptr;
  const bool on = (c->timing >> phase) & 1u;
  if (on) {
    a = pool_event(c);
    b = pool_event(c);
    if (!a || !b) return fail(GKS_ERR_HIP, "hipEventCreate");
    HIPCHK(hipEventRecord(a, c->s));
  }
  HIPCHK(launch());
  if (on) {
    HIPCHK(hipEventRecord(b, c->s));
    c->pending.push_back({phase, a, b});
  }
  return GKS_OK;
}
// folds finished event pairs into the accumulators (stream must be idle)
static int collect_timing(gks3d_ctx* c) {
  for (auto& r : c->pending) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    c->t_ms[r.phase] += ms;
    c->t_n[r.phase]++;
  }
  c->pending.clear();
  c->pool_used = 0;
  return GKS_OK;
}

static int dev_alloc(gks3d_ctx* c, double** out, size_t n) {
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, n * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  c->allocs.push_back(ptr);
  e = hipMemsetAsync(ptr, 0, n * sizeof(double), c->s);
  if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
  *out = static_cast<double*>(ptr);
  return GKS_OK;
}

// Get_CFL / Dtx_3d (time_advance.cpp:793-848) from the largest |u| + c and
// the smallest density of the last primitive sweep: dt starts at the
// smallest cell size (R_c), the acoustic limit CFL h / (|u| + c), the
// viscous limit 0.33 CFL h^2 rho / Mu for Navier-Stokes, then the clamp to
// tstop
static double cfl_dt(const gks3d_ctx* c, double t, double tstop) {
  const Dev3& d = c->d;
  double h = d.dx;
  if (d.dy < h) h = d.dy;
  if (d.dz < h) h = d.dz;
  const double cfl = c->p.cfl;
  double dt = h;
  if (c->tmax > cfl * h / dt) dt = cfl * h / c->tmax;
  if (dt > 0.33 * cfl * h * h * c->rhomin / c->p.mu && c->p.tau_type != GKS_TAU_EULER && c->p.mu > 0)
    dt = 0.33 * cfl * h * h * c->rhomin / c->p.mu;
  if (t + dt > tstop) dt = tstop - t + 1e-16;
  return dt;
}

extern "C" {

const char* gks3d_last_error(void) { return g_err3.c_str(); }

int gks3d_destroy(gks3d_ctx* c) {
  if (!c) return GKS_OK;
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (void* a : c->allocs) (void)hipFree(a);
  if (c->d_mlp) (void)hipFree(c->d_mlp);
  if (c->d_mlpw) (void)hipFree(c->d_mlpw);
  if (c->d_lin) (void)hipFree(c->d_lin);
  if (c->d_lstat) (void)hipFree(c->d_lstat);
  if (c->h_stat) (void)hipHostFree(c->h_stat);
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return GKS_OK;
}

int gks3d_create(const gks3d_params* p, gks3d_ctx** out) {
  if (!out) return fail(GKS_ERR_ARG, "null out");
  *out = nullptr;
  RC(check_params(p));
  gks3d_ctx* c = new gks3d_ctx();
  c->p = *p;
  Dev3& d = c->d;
  std::memset(&d, 0, sizeof(d));
  d.NX = p->nodex;
  d.NY = p->nodey;
  d.NZ = p->nodez;
  d.G = p->ghost;
  d.nx = d.NX + 2 * d.G;
  d.ny = d.NY + 2 * d.G;
  d.nz = d.NZ + 2 * d.G;
  d.nc = (long long)d.nx * d.ny * d.nz;
  d.nf = (long long)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
  d.recon_var = p->recon_variable;
  // SetUniformMesh (3D): basic_function.cpp via 3Dproblem.cpp:84-86
  d.dx = (p->right - p->left) / p->nodex;
  d.dy = (p->up - p->down) / p->nodey;
  d.dz = (p->front - p->back) / p->nodez;
  d.gas = gas_of(*p);
  c->lt = p->fast_math ? &fast::table3d : &ieee::table3d;
  hipError_t e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c->h_stat), 4 * sizeof(double));
  if (e != hipSuccess) {
    gks3d_destroy(c);
    return hip_fail(e, "gks3d_create");
  }
  const size_t nc = d.nc, nf = d.nf;
  int rc;
  if ((rc = dev_alloc(c, &d.U, 5 * nc)) || (rc = dev_alloc(c, &d.Uold, 5 * nc)) ||
      (rc = dev_alloc(c, &d.fv, 20 * nf)) || (rc = dev_alloc(c, &d.lv, 60 * nf)) ||
      (rc = dev_alloc(c, &d.part, 3 * kParts3)) || (rc = dev_alloc(c, &d.stat, 4))) {
    gks3d_destroy(c);
    return rc;
  }
  for (int dir = 0; dir < 3; dir++)
    if ((rc = dev_alloc(c, &d.xs[dir], 5 * nf)) || (rc = dev_alloc(c, &d.pc[dir], 5 * nf))) {
      gks3d_destroy(c);
      return rc;
    }
  e = hipStreamSynchronize(c->s);
  if (e != hipSuccess) {
    gks3d_destroy(c);
    return hip_fail(e, "gks3d_create");
  }
  *out = c;
  return GKS_OK;
}

static void gather(std::vector<double>& soa, const double* base, long long n, int64_t cs, int64_t vs) {
  const char* b = reinterpret_cast<const char*>(base);
  for (long long k = 0; k < n; k++)
    for (int v = 0; v < 5; v++) soa[v * n + k] = *reinterpret_cast<const double*>(b + k * cs + v * vs);
}

int gks3d_set_convar(gks3d_ctx* c, const double* base, int64_t cs, int64_t vs) {
  if (!c || !base) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(5 * (size_t)c->d.nc);
  gather(h, base, c->d.nc, cs, vs);
  HIPCHK(hipMemcpyAsync(c->d.U, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return GKS_OK;
}
int gks3d_get_convar(gks3d_ctx* c, double* base, int64_t cs, int64_t vs) {
  if (!c || !base) return fail(GKS_ERR_ARG, "null");
  const long long n = c->d.nc;
  std::vector<double> h(5 * (size_t)n);
  HIPCHK(hipMemcpyAsync(h.data(), c->d.U, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  char* b = reinterpret_cast<char*>(base);
  for (long long k = 0; k < n; k++)
    for (int v = 0; v < 5; v++) *reinterpret_cast<double*>(b + k * cs + v * vs) = h[v * n + k];
  return GKS_OK;
}
// the CFL inputs of the first step from the given primitive variables
// (Get_CFL reads primvar, which at step 0 holds the initial condition's)
int gks3d_set_cfl_primvar(gks3d_ctx* c, const double* base, int64_t cs, int64_t vs) {
  if (!c || !base) return fail(GKS_ERR_ARG, "null");
  const Dev3& d = c->d;
  const char* b = reinterpret_cast<const char*>(base);
  double tmax = 0.0, rmin = 1e300;
  for (int i = d.G; i < d.G + d.NX; i++)
    for (int j = d.G; j < d.G + d.NY; j++)
      for (int k = d.G; k < d.G + d.NZ; k++) {
        const long long cell = ((long long)i * d.ny + j) * d.nz + k;
        double p[5];
        for (int v = 0; v < 5; v++) p[v] = *reinterpret_cast<const double*>(b + cell * cs + v * vs);
        const double overden = 1.0 / p[0];
        const double s = std::sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) + std::sqrt(d.gas.gamma * p[4] * overden);
        if (s > tmax) tmax = s;
        if (p[0] < rmin) rmin = p[0];
      }
  c->tmax = tmax;
  c->rhomin = rmin;
  c->have_speed = true;
  return GKS_OK;
}

static int stage(gks3d_ctx* c, int s, double dt, const int* bc) {
  const Dev3& d = c->d;
  const Launch3dTable* lt = c->lt;
  RC(timed(c, GKS3_PHASE_BC, [&] { return lt->bc(d, bc, c->s); }));
  RC(timed(c, GKS3_PHASE_PRIMVAR, [&] { return lt->primvar(d, c->s); }));
  if (d.mlp) RC(timed(c, GKS3_PHASE_CLASSIFY, [&] { return lt->classify(d, c->s); }));
  for (int dir = 0; dir < 3; dir++) {
    RC(timed(c, GKS3_PHASE_FACE, [&] { return lt->face(d, dir, c->s); }));
    RC(timed(c, GKS3_PHASE_LINE, [&] { return lt->line(d, dir, c->s); }));
    RC(timed(c, GKS3_PHASE_FLUX, [&] { return lt->flux(d, dir, s, dt, c->s); }));
  }
  RC(timed(c, GKS3_PHASE_UPDATE, [&] { return lt->update(d, c->s); }));
  return GKS_OK;
}

int gks3d_run(gks3d_ctx* c, const int32_t bc[6], double tstop, int32_t max_steps, double* t_inout,
              int32_t* step_inout) {
  if (!c || !bc || !t_inout || !step_inout) return fail(GKS_ERR_ARG, "null");
  if (!c->have_speed) return fail(GKS_ERR_STATE, "no primitive state yet: call gks3d_set_cfl_primvar first");
  const Dev3& d = c->d;
  int b[6];
  for (int s = 0; s < 6; s++) {
    if (bc[s] != GKS_BC_FREE && bc[s] != GKS_BC_PERIODIC) return fail(GKS_ERR_UNSUPPORTED, "3D boundaries are free or periodic");
    b[s] = bc[s];
  }
  for (int a = 0; a < 3; a++) {
    const bool per_lo = b[2 * a] == GKS_BC_PERIODIC, per_hi = b[2 * a + 1] == GKS_BC_PERIODIC;
    const int N = a == 0 ? d.NX : (a == 1 ? d.NY : d.NZ);
    if ((per_lo || per_hi) && N < d.G) return fail(GKS_ERR_ARG, "periodic boundaries need at least ghost interior cells");
  }
  if (max_steps < 0) return fail(GKS_ERR_ARG, "max_steps < 0");
  double t = *t_inout;
  int32_t step = *step_inout;
  int n = 0;
  bool blew = false;
  while (t < tstop && n < max_steps) {
    const double dt = cfl_dt(c, t, tstop);
    RC(timed(c, GKS3_PHASE_COPY_OLD, [&] { return c->lt->copy_old(d, c->s); }));
    RC(stage(c, 0, dt, b));
    RC(stage(c, 1, dt, b));
    HIPCHK(hipMemcpyAsync(c->h_stat, d.stat, 4 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->stat_pending = false;
    if (c->timing) RC(collect_timing(c));
    if (c->h_stat[2] != 0.0) {
      blew = true;
      break;
    }
    c->tmax = c->h_stat[0];
    c->rhomin = c->h_stat[1];
    t = t + dt;
    step++;
    n++;
  }
  *t_inout = t;
  *step_inout = step;
  if (blew) {
    const double zero = 0.0;
    HIPCHK(hipMemcpyAsync(d.stat + 2, &zero, sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return fail(GKS_ERR_BLOWUP, "density or pressure negative or NaN");
  }
  return GKS_OK;
}

int gks3d_boundary(gks3d_ctx* c, int32_t side, int32_t kind) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (side < 0 || side > 5) return fail(GKS_ERR_ARG, "side");
  if (kind != GKS_BC_FREE && kind != GKS_BC_PERIODIC) return fail(GKS_ERR_UNSUPPORTED, "3D boundaries are free or periodic");
  const Dev3& d = c->d;
  const int N = side < 2 ? d.NX : (side < 4 ? d.NY : d.NZ);
  if (kind == GKS_BC_PERIODIC && N < d.G) return fail(GKS_ERR_ARG, "periodic boundaries need at least ghost interior cells");
  return timed(c, GKS3_PHASE_BC, [&] { return c->lt->bc_side(d, side, kind, c->s); });
}
int gks3d_primvar(gks3d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  RC(timed(c, GKS3_PHASE_PRIMVAR, [&] { return c->lt->primvar(c->d, c->s); }));
  c->stat_pending = true;
  return GKS_OK;
}
int gks3d_copy_old(gks3d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  return timed(c, GKS3_PHASE_COPY_OLD, [&] { return c->lt->copy_old(c->d, c->s); });
}
int gks3d_cfl(gks3d_ctx* c, double t, double tstop, double* dt) {
  if (!c || !dt) return fail(GKS_ERR_ARG, "null");
  if (c->stat_pending) {
    HIPCHK(hipMemcpyAsync(c->h_stat, c->d.stat, 4 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (c->timing) RC(collect_timing(c));
    c->stat_pending = false;
    if (c->h_stat[2] != 0.0) {
      const double zero = 0.0;
      HIPCHK(hipMemcpyAsync(c->d.stat + 2, &zero, sizeof(double), hipMemcpyHostToDevice, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
      return fail(GKS_ERR_BLOWUP, "density or pressure negative or NaN");
    }
    c->tmax = c->h_stat[0];
    c->rhomin = c->h_stat[1];
    c->have_speed = true;
  }
  if (!c->have_speed) return fail(GKS_ERR_STATE, "no primitive state yet: call gks3d_set_cfl_primvar first");
  *dt = cfl_dt(c, t, tstop);
  return GKS_OK;
}
int gks3d_flux(gks3d_ctx* c, int32_t stage, double dt) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (stage < 0 || stage > 1) return fail(GKS_ERR_ARG, "stage");
  const Dev3& d = c->d;
  const Launch3dTable* lt = c->lt;
  if (d.mlp) RC(timed(c, GKS3_PHASE_CLASSIFY, [&] { return lt->classify(d, c->s); }));
  for (int dir = 0; dir < 3; dir++) {
    RC(timed(c, GKS3_PHASE_FACE, [&] { return lt->face(d, dir, c->s); }));
    RC(timed(c, GKS3_PHASE_LINE, [&] { return lt->line(d, dir, c->s); }));
    RC(timed(c, GKS3_PHASE_FLUX, [&] { return lt->flux(d, dir, stage, dt, c->s); }));
  }
  return GKS_OK;
}
int gks3d_update(gks3d_ctx* c, int32_t stage) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (stage < 0 || stage > 1) return fail(GKS_ERR_ARG, "stage");
  return timed(c, GKS3_PHASE_UPDATE, [&] { return c->lt->update(c->d, c->s); });
}

int gks3d_timing_phases(gks3d_ctx* c, uint32_t mask) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(hipStreamSynchronize(c->s));
  RC(collect_timing(c));
  for (int k = 0; k < GKS3_NPHASE; k++) {
    c->t_ms[k] = 0.0;
    c->t_n[k] = 0;
  }
  c->timing = mask & ((1u << GKS3_NPHASE) - 1u);
  return GKS_OK;
}
int gks3d_timing(gks3d_ctx* c, int enable) { return gks3d_timing_phases(c, enable ? (1u << GKS3_NPHASE) - 1u : 0u); }
int gks3d_get_timing(gks3d_ctx* c, double* ms, int64_t* launches) {
  if (!c || !ms || !launches) return fail(GKS_ERR_ARG, "null");
  HIPCHK(hipStreamSynchronize(c->s));
  RC(collect_timing(c));
  for (int k = 0; k < GKS3_NPHASE; k++) {
    ms[k] = c->t_ms[k];
    launches[k] = c->t_n[k];
  }
  return GKS_OK;
}

int gks3d_sync(gks3d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(hipStreamSynchronize(c->s));
  return GKS_OK;
}

constexpr size_t kLstat3Words = 3 * kStat3Slots * 2;
int gks3d_set_classifier(gks3d_ctx* c, const float* params, float logit_threshold) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(hipStreamSynchronize(c->s));
  if (!params) {
    c->d.mlp = nullptr;
    return GKS_OK;
  }
  Dev3& d = c->d;
  if (!c->d_mlp) {
    HIPCHK(hipMalloc(&c->d_mlp, GKS_MLP_NPARAM * sizeof(float)));
    HIPCHK(hipMalloc(&c->d_mlpw, mlp::kPackHalves * sizeof(unsigned short)));
    HIPCHK(hipMalloc(&c->d_lin, 3 * (size_t)d.nc + 2 * (size_t)d.nf));
    HIPCHK(hipMalloc(&c->d_lstat, kLstat3Words * sizeof(unsigned long long)));
  }
  std::vector<uint16_t> packed(mlp::kPackHalves);
  mlp::pack_bf16(params, packed.data());
  HIPCHK(hipMemcpy(c->d_mlp, params, GKS_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_mlpw, packed.data(), packed.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->d_lin, 0, 3 * (size_t)d.nc + 2 * (size_t)d.nf));
  HIPCHK(hipMemset(c->d_lstat, 0, kLstat3Words * sizeof(unsigned long long)));
  d.mlp = c->d_mlp;
  d.mlpw = c->d_mlpw;
  d.mlp_thr = logit_threshold;
  d.lin3 = c->d_lin;
  d.lin_line = c->d_lin + 3 * (size_t)d.nc;
  d.lstat3 = c->d_lstat;
  return GKS_OK;
}

int gks3d_classifier_counts(gks3d_ctx* c, int64_t counts[6]) {
  if (!c || !counts) return fail(GKS_ERR_ARG, "null");
  if (!c->d_lstat) return fail(GKS_ERR_STATE, "no classifier set");
  std::vector<unsigned long long> h(kLstat3Words);
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(h.data(), c->d_lstat, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int k = 0; k < 6; k++) counts[k] = 0;
  for (int kind = 0; kind < 3; kind++)
    for (int q = 0; q < kStat3Slots; q++) {
      counts[2 * kind] += (int64_t)h[2 * (kind * kStat3Slots + q)];
      counts[2 * kind + 1] += (int64_t)h[2 * (kind * kStat3Slots + q) + 1];
    }
  return GKS_OK;
}

int gks3d_flux_points(const gks3d_params* p, int64_t n, const double* left, const double* right, double dt,
                      double* center, double* f, double* derf) {
  if (!p || !left || !right || !f || !derf) return fail(GKS_ERR_ARG, "null");
  if (n < 0) return fail(GKS_ERR_ARG, "n < 0");
  if (n == 0) return GKS_OK;
  const Launch3dTable* lt = p->fast_math ? &fast::table3d : &ieee::table3d;
  double *dl = nullptr, *dr = nullptr, *dc = nullptr, *dfo = nullptr, *dd = nullptr;
  hipError_t e = hipMalloc(&dl, n * 20 * 8);
  if (e == hipSuccess) e = hipMalloc(&dr, n * 20 * 8);
  if (e == hipSuccess) e = hipMalloc(&dc, n * 20 * 8);
  if (e == hipSuccess) e = hipMalloc(&dfo, n * 5 * 8);
  if (e == hipSuccess) e = hipMalloc(&dd, n * 5 * 8);
  if (e == hipSuccess) e = hipMemcpy(dl, left, n * 20 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dr, right, n * 20 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = lt->flux_points(gas_of(*p), n, dl, dr, dt, center ? dc : nullptr, dfo, dd, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && center) e = hipMemcpy(center, dc, n * 20 * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(f, dfo, n * 5 * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(derf, dd, n * 5 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(dl);
  (void)hipFree(dr);
  (void)hipFree(dc);
  (void)hipFree(dfo);
  (void)hipFree(dd);
  if (e != hipSuccess) return hip_fail(e, "gks3d_flux_points");
  return GKS_OK;
}

}  // extern "C"
