// gks2d_api.cpp — the extern "C" layer of include/gks2d.h over the HIP
// kernels.  One context = one block = one HIP stream; a context is not
// re-entrant (the reference is single-threaded at this level too).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gks2d.h"
#include "../../include/gks_learn.h"
#include "gks2d_launch.hpp"
#include "mlp_pack.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* what) {
  g_err = what;
  return code;
}
int hip_fail(hipError_t e, const char* where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return GKS_ERR_HIP;
}
#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

gksd::Gas gas_of(const gks2d_params& p) {
  gksd::Gas g;
  g.K = p.K;
  g.gamma = p.gamma;
  g.mu = p.mu;
  g.nu = p.nu;
  g.pr = p.pr;
  g.c1 = p.c1_euler;
  g.c2 = p.c2_euler;
  g.tau_type = p.tau_type;
  g.prandtl_fix = p.prandtl_fix;
  g.inv_gm1 = 1.0 / (p.gamma - 1.0);
  return g;
}

int check_params(const gks2d_params* p) {
  if (!p) return fail(GKS_ERR_ARG, "null params");
  if (p->ghost != 5) return fail(GKS_ERR_UNSUPPORTED, "ghost must be 5");
  if (p->nodex < 5 || p->nodey < 5) return fail(GKS_ERR_ARG, "nodex and nodey must be >= 5");
  if (!((p->weno_order == 5 && p->gausspoint == 2) || (p->weno_order == 7 && p->gausspoint == 3) ||
        (p->weno_order == 9 && p->gausspoint == 4)))
    return fail(GKS_ERR_UNSUPPORTED, "implemented: WENO5/7/9-AO with 2/3/4 Gauss points");
  if (p->use_df != 1) return fail(GKS_ERR_UNSUPPORTED, "implemented: the DF (discontinuity feedback) variants");
  if (p->recon_variable != GKS_VAR_CONSERVATIVE && p->recon_variable != GKS_VAR_CHARACTERISTIC)
    return fail(GKS_ERR_ARG, "recon_variable");
  if (p->right <= p->left || p->up <= p->down) return fail(GKS_ERR_ARG, "empty domain");
  if (!(p->gamma > 1.0) || p->K < 0) return fail(GKS_ERR_ARG, "gamma / K");
  return GKS_OK;
}

}  // namespace

struct gks2d_ctx {
  gks2d_params p;
  Dev d;
  const LaunchTable* lt = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  double* h_stat = nullptr;  // pinned copy of d.stat
  double* d_prim = nullptr;  // staging for gks2d_set_cfl_primvar
  float* d_mlp = nullptr;    // classifier parameters (gks2d_set_classifier)
  unsigned short* d_mlpw = nullptr;  // their weights in the MFMA operand layout
  unsigned char* d_lin = nullptr;    // per-cell linear-weight flags, x then y
  unsigned* d_tlist = nullptr;       // learned path: listed x, y faces; listed x, y normal cells; 4 counts
  unsigned long long* d_lstat = nullptr;  // its linear / total counters
  std::vector<void*> allocs;
  int last_slot = -1;        // slot of the last primvar / set_cfl_primvar
  int last_flux_stage = -1;
  int kept_stage_mask = 0;   // stages whose raw face fluxes are stored
  // per-phase HIP-event timing (gks2d_timing)
  unsigned timing = 0;  // bit mask of the timed phases
  struct Rec {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  double t_ms[PH_COUNT] = {};
  long long t_n[PH_COUNT] = {};
};

static hipEvent_t pool_event(gks2d_ctx* c) {
  if (c->pool_used == c->pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->pool.push_back(e);
  }
  return c->pool[c->pool_used++];
}
// launches one kernel phase; with timing on, brackets it with events on the
// context's stream
static int collect_timing(gks2d_ctx* c);
// Timing events are read back in batches of this many timed launches (which
// bounds the event pool): draining the stream after every step would keep
// the next step's launches from queueing behind the current one.
constexpr size_t kTimingBatch = 2048;
template <class F> static int timed(gks2d_ctx* c, int phase, F&& launch) {
  hipEvent_t a = nullptr, b = nullptr;
  const bool on = (c->timing >> phase) & 1u;
  if (on && c->pending.size() >= kTimingBatch) {
    HIPCHK(hipStreamSynchronize(c->s));
    const int rc = collect_timing(c);
    if (rc != GKS_OK) return rc;
  }
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
static int collect_timing(gks2d_ctx* c) {
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
#define RC(x)              \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

static int dev_alloc(gks2d_ctx* c, double** out, size_t n) {
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

const char* gks2d_last_error(void) { return g_err.c_str(); }

int gks2d_destroy(gks2d_ctx* c) {
  if (!c) return GKS_OK;
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (void* a : c->allocs) (void)hipFree(a);
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  if (c->h_stat) (void)hipHostFree(c->h_stat);
  if (c->d_mlp) (void)hipFree(c->d_mlp);
  if (c->d_mlpw) (void)hipFree(c->d_mlpw);
  if (c->d_lin) (void)hipFree(c->d_lin);
  if (c->d_tlist) (void)hipFree(c->d_tlist);
  if (c->d_lstat) (void)hipFree(c->d_lstat);
  if (c->ev) (void)hipEventDestroy(c->ev);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return GKS_OK;
}

int gks2d_create(const gks2d_params* p, gks2d_ctx** out) {
  if (!out) return fail(GKS_ERR_ARG, "null out");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  gks2d_ctx* c = new gks2d_ctx();
  c->p = *p;
  Dev& d = c->d;
  std::memset(&d, 0, sizeof(d));
  d.NX = p->nodex;
  d.NY = p->nodey;
  d.G = p->ghost;
  d.nx = d.NX + 2 * d.G;
  d.ny = d.NY + 2 * d.G;
  d.gp = p->gausspoint;
  d.recon_var = p->recon_variable;
  d.use_df = p->use_df;
  d.dx = (p->right - p->left) / p->nodex;  // SetUniformMesh, basic_function.cpp:1411-1416
  d.dy = (p->up - p->down) / p->nodey;
  d.area = d.dx * d.dy;
  d.idx = 1.0 / d.dx;
  d.idy = 1.0 / d.dy;
  d.order = p->weno_order;
  // the tangential reconstruction runs inside the flux launch (its
  // Gauss-point states stay in LDS); GKS_TANFLUX=0 keeps the two launches
  d.df_frozen = p->df_frozen ? 1 : 0;
  d.fuse_tanflux = !(std::getenv("GKS_TANFLUX") && std::getenv("GKS_TANFLUX")[0] == '0');
  // the two Gauss points of a face meet in one lane pair of the fused
  // launch, which stores per-face values for k_alpha and k_update
  d.presum = d.fuse_tanflux && p->weno_order == 5;
  if (d.gp == 2) {  // SetGuassPoint, fluid_mesh.cpp:9-45
    d.gw[0] = 0.5;
    d.gw[1] = 0.5;
  } else if (d.gp == 3) {
    d.gw[0] = 4.0 / 9.0;
    d.gw[1] = 0.5 - 0.5 * d.gw[0];
    d.gw[2] = d.gw[1];
  } else {
    d.gw[0] = 1.0 / 12.0;
    d.gw[1] = 1.0 / 12.0;
    d.gw[2] = 5.0 / 12.0;
    d.gw[3] = 5.0 / 12.0;
  }
  d.gas = gas_of(*p);
  d.nc = (long long)d.nx * d.ny;
  d.nfx = (long long)(d.nx + 1) * d.ny;
  d.nfy = (long long)d.nx * (d.ny + 1);
  c->lt = p->fast_math ? &fast::table : &ieee::table;
  const long long nfm = d.nfx > d.nfy ? d.nfx : d.nfy;
  (void)nfm;
  hipError_t e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    gks2d_destroy(c);
    return hip_fail(e, "hipStreamCreate");
  }
  e = hipEventCreateWithFlags(&c->ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_stat, 8 * sizeof(double));
  if (e != hipSuccess) {
    gks2d_destroy(c);
    return hip_fail(e, "hipEventCreate/hipHostMalloc");
  }
  std::memset(c->h_stat, 0, 8 * sizeof(double));
  const int gp = d.gp;
  if ((rc = dev_alloc(c, &d.U, 4 * d.nc)) || (rc = dev_alloc(c, &d.Uold, 4 * d.nc)) ||
      (rc = dev_alloc(c, &d.ax, d.nc)) || (rc = dev_alloc(c, &d.ay, d.nc)) ||
      (rc = dev_alloc(c, &d.line_x, 16 * d.nfx)) || (rc = dev_alloc(c, &d.line_y, 16 * d.nfy)) ||
      // Gauss-point states reach HBM only when the tangential launch is separate
      (!d.fuse_tanflux && ((rc = dev_alloc(c, &d.gst_x, 24 * gp * d.nfx)) ||
                           (rc = dev_alloc(c, &d.gst_y, 24 * gp * d.nfy)))) ||
      (rc = dev_alloc(c, &d.xcx, 4 * gp * d.nfx)) || (rc = dev_alloc(c, &d.xcy, 4 * gp * d.nfy)) ||
      (rc = dev_alloc(c, &d.pcx, 4 * gp * d.nfx)) || (rc = dev_alloc(c, &d.pcy, 4 * gp * d.nfy)) ||
      (rc = dev_alloc(c, &d.akx, gp * d.nfx)) || (rc = dev_alloc(c, &d.aky, gp * d.nfy)) ||
      (rc = dev_alloc(c, &d.partial, 2 * std::max<long long>(kSweepPartials, (d.nc + kSweepBlock - 1) / kSweepBlock))) ||
      (rc = dev_alloc(c, &d.stat, 8)) ||
      (rc = dev_alloc(c, &c->d_prim, 4 * d.nc))) {
    gks2d_destroy(c);
    return rc;
  }
  *out = c;
  return GKS_OK;
}

// ------------------------------------------------------------ host copies
static void gather(std::vector<double>& soa, const double* base, long long nc, int nvar, int64_t cs, int64_t vs) {
  const char* b = reinterpret_cast<const char*>(base);
  for (long long k = 0; k < nc; k++)
    for (int v = 0; v < nvar; v++) soa[v * nc + k] = *reinterpret_cast<const double*>(b + k * cs + v * vs);
}
static void scatter(const std::vector<double>& soa, double* base, long long nc, int nvar, int64_t cs, int64_t vs) {
  char* b = reinterpret_cast<char*>(base);
  for (long long k = 0; k < nc; k++)
    for (int v = 0; v < nvar; v++) *reinterpret_cast<double*>(b + k * cs + v * vs) = soa[v * nc + k];
}

int gks2d_set_convar(gks2d_ctx* c, const double* base, int64_t cs, int64_t vs) {
  if (!c || !base) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(4 * c->d.nc);
  gather(h, base, c->d.nc, 4, cs, vs);
  HIPCHK(hipMemcpyAsync(c->d.U, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return GKS_OK;
}
int gks2d_get_convar(gks2d_ctx* c, double* base, int64_t cs, int64_t vs) {
  if (!c || !base) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(4 * c->d.nc);
  HIPCHK(hipMemcpyAsync(h.data(), c->d.U, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  scatter(h, base, c->d.nc, 4, cs, vs);
  return GKS_OK;
}
int gks2d_set_alpha(gks2d_ctx* c, const double* ax, const double* ay, int64_t cs) {
  if (!c || !ax || !ay) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(c->d.nc);
  gather(h, ax, c->d.nc, 1, cs, 0);
  HIPCHK(hipMemcpyAsync(c->d.ax, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  gather(h, ay, c->d.nc, 1, cs, 0);
  HIPCHK(hipMemcpyAsync(c->d.ay, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return GKS_OK;
}
int gks2d_get_alpha(gks2d_ctx* c, double* ax, double* ay, int64_t cs) {
  if (!c || !ax || !ay) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(c->d.nc);
  HIPCHK(hipMemcpyAsync(h.data(), c->d.ax, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  scatter(h, ax, c->d.nc, 1, cs, 0);
  HIPCHK(hipMemcpyAsync(h.data(), c->d.ay, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  scatter(h, ay, c->d.nc, 1, cs, 0);
  return GKS_OK;
}
int gks2d_set_cfl_primvar(gks2d_ctx* c, const double* base, int64_t cs, int64_t vs) {
  if (!c || !base) return fail(GKS_ERR_ARG, "null");
  std::vector<double> h(4 * c->d.nc);
  gather(h, base, c->d.nc, 4, cs, vs);
  HIPCHK(hipMemcpyAsync(c->d_prim, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->s));
  HIPCHK(c->lt->primvar(c->d, c->d_prim, 0, 4, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  c->last_slot = 0;
  return GKS_OK;
}

// ------------------------------------------------------------ sweeps
int gks2d_copy_old(gks2d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  return timed(c, PH_COPY_OLD, [&] { return c->lt->copy_old(c->d, c->s); });
}
int gks2d_boundary(gks2d_ctx* c, int side, int kind) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (side < 0 || side > 3 || kind < 0 || kind > 2) return fail(GKS_ERR_ARG, "side/kind");
  return timed(c, PH_BC, [&] { return c->lt->bc(c->d, side, kind, c->s); });
}
int gks2d_primvar(gks2d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  RC(timed(c, PH_PRIMVAR, [&] { return c->lt->primvar(c->d, nullptr, 0, 4, c->s); }));
  c->last_slot = 0;
  return GKS_OK;
}
// learned path whose smooth-face launches form every line value (one
// classifier launch for both directions): each kind's two lists run in one
// launch
static bool both_lists(const Dev& d) {
  return d.mlp && d.smooth_flux && d.order == 5 && d.fuse_tanflux && (d.fused_lines & 3) == 3;
}
int gks2d_reconstruct(gks2d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  RC(timed(c, PH_NORMAL_X, [&] { return c->lt->normal(c->d, 0, c->s); }));
  if (both_lists(c->d)) {
    // both directions' classifiers ran in the x call; their listed cells
    // follow in one launch (timed as the x phase)
    RC(timed(c, PH_NORMAL_LIST_X, [&] { return c->lt->normal_list(c->d, kBothDirs, c->s); }));
  } else {
    if (c->d.mlp) RC(timed(c, PH_NORMAL_LIST_X, [&] { return c->lt->normal_list(c->d, 0, c->s); }));
    RC(timed(c, PH_NORMAL_Y, [&] { return c->lt->normal(c->d, 1, c->s); }));
    if (c->d.mlp) RC(timed(c, PH_NORMAL_LIST_Y, [&] { return c->lt->normal_list(c->d, 1, c->s); }));
  }
  RC(timed(c, PH_TANGENT_X, [&] { return c->lt->tangent(c->d, 0, c->s); }));
  return timed(c, PH_TANGENT_Y, [&] { return c->lt->tangent(c->d, 1, c->s); });
}
int gks2d_flux(gks2d_ctx* c, int stage, double dt) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (stage < 0 || stage > 1) return fail(GKS_ERR_ARG, "stage");
  const bool listed = c->d.mlp && c->d.smooth_flux;
  if (both_lists(c->d)) {
    // both smooth-face launches, then both directions' listed faces (one
    // prefix launch; timed as the x phase): no list reads what the other
    // direction's smooth-face launch writes
    RC(timed(c, PH_FLUX_X, [&] { return c->lt->flux(c->d, 0, stage, dt, c->s); }));
    RC(timed(c, PH_FLUX_Y, [&] { return c->lt->flux(c->d, 1, stage, dt, c->s); }));
    RC(timed(c, PH_FLUX_LIST_X, [&] { return c->lt->flux_list(c->d, kBothDirs, stage, dt, c->s); }));
  } else {
    RC(timed(c, PH_FLUX_X, [&] { return c->lt->flux(c->d, 0, stage, dt, c->s); }));
    if (listed) RC(timed(c, PH_FLUX_LIST_X, [&] { return c->lt->flux_list(c->d, 0, stage, dt, c->s); }));
    RC(timed(c, PH_FLUX_Y, [&] { return c->lt->flux(c->d, 1, stage, dt, c->s); }));
    if (listed) RC(timed(c, PH_FLUX_LIST_Y, [&] { return c->lt->flux_list(c->d, 1, stage, dt, c->s); }));
  }
  c->last_flux_stage = stage;
  if (c->d.keep_face_fluxes) c->kept_stage_mask |= 1 << stage;
  return GKS_OK;
}
int gks2d_update(gks2d_ctx* c, int stage) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (stage < 0 || stage > 1) return fail(GKS_ERR_ARG, "stage");
  return timed(c, PH_UPDATE, [&] { return c->lt->update(c->d, stage, c->s); });
}
int gks2d_update_alpha(gks2d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  return timed(c, PH_ALPHA, [&] { return c->lt->alpha(c->d, c->s); });
}
int gks2d_timing_phases(gks2d_ctx* c, uint32_t mask) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(hipStreamSynchronize(c->s));
  RC(collect_timing(c));
  for (int k = 0; k < PH_COUNT; k++) {
    c->t_ms[k] = 0.0;
    c->t_n[k] = 0;
  }
  c->timing = mask & ((1u << PH_COUNT) - 1u);
  return GKS_OK;
}
int gks2d_timing(gks2d_ctx* c, int enable) { return gks2d_timing_phases(c, enable ? (1u << PH_COUNT) - 1u : 0u); }
int gks2d_get_timing(gks2d_ctx* c, double* ms, int64_t* launches) {
  if (!c || !ms || !launches) return fail(GKS_ERR_ARG, "null");
  HIPCHK(hipStreamSynchronize(c->s));
  RC(collect_timing(c));
  for (int k = 0; k < PH_COUNT; k++) {
    ms[k] = c->t_ms[k];
    launches[k] = c->t_n[k];
  }
  return GKS_OK;
}
int gks2d_sync(gks2d_ctx* c) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(hipStreamSynchronize(c->s));
  return GKS_OK;
}

// Get_CFL / Dtx (time_advance.cpp:264-305) from the max speed over the
// interior.  Equals the reference's sequential scan except when two cells'
// speeds differ by one ulp (the scan may then keep the earlier one).
static double dt_from_speed(const gks2d_ctx* c, double tmax, double t, double tstop) {
  const gks2d_params& p = c->p;
  double dt = c->d.dx;
  const double hs[2] = {c->d.dx, c->d.dy};
  for (double h : hs) {
    if (tmax > p.cfl * h / dt) dt = p.cfl * h / tmax;
    if (dt > 0.25 * p.cfl * h * h / p.mu && p.tau_type == GKS_TAU_NS && p.mu > 0) dt = 0.25 * p.cfl * h * h / p.mu;
  }
  if (t + dt > tstop) dt = tstop - t + 1e-16;
  return dt;
}

int gks2d_cfl(gks2d_ctx* c, double t, double tstop, double* dt_out) {
  if (!c || !dt_out) return fail(GKS_ERR_ARG, "null");
  if (c->last_slot < 0) return fail(GKS_ERR_STATE, "no primitive state yet: call gks2d_set_cfl_primvar first");
  HIPCHK(hipMemcpyAsync(c->h_stat, c->d.stat, 8 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (c->h_stat[4] != 0.0) {
    const double zero = 0.0;
    HIPCHK(hipMemcpyAsync(c->d.stat + 4, &zero, sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return fail(GKS_ERR_BLOWUP, "density or pressure negative or NaN");
  }
  *dt_out = dt_from_speed(c, c->h_stat[2 * c->last_slot], t, tstop);
  return GKS_OK;
}

// one stage of the driver loop (2Dproblem.cpp:137-158)
static int stage_launch(gks2d_ctx* c, int stage, double dt, const int32_t* bc) {
  const int kinds[4] = {bc[0], bc[1], bc[2], bc[3]};
  RC(timed(c, PH_BC, [&] { return c->lt->bc_all(c->d, kinds, c->s); }));
  RC(timed(c, PH_PRIMVAR, [&] { return c->lt->primvar(c->d, nullptr, stage, 4, c->s); }));
  RC(gks2d_reconstruct(c));
  RC(gks2d_flux(c, stage, dt));
  RC(gks2d_update(c, stage));
  if (c->p.use_df && !c->p.df_frozen) RC(gks2d_update_alpha(c));
  return GKS_OK;
}

// One S2O4 step of gks2d_run without CopyFluid_new_to_old: with A the
// state (d.U) and B the second buffer (d.Uold), stage 0 reconstructs A and
// writes its state A + dt F0 to B; stage 1 reconstructs B and updates A in
// place (each cell reads its A value before writing it).  The arithmetic
// is that of copy + two stages; the step ends with the state in A, whose
// ghost frame still holds stage 0's fill (stage 1 filled B's): the run
// copies B's frame into A when it ends, as the state of copy + two stages
// has it.
//
// Convar_to_Primvar's blow-up test and CFL speed come from the updates
// (update_stat): stage 0's, on the state stage 1 starts from, into slot 1
// (the next step's stale CFL input, read back right away), stage 1's, on
// the state the next step starts from, into slot 0.  Only the run's first
// stage runs the primvar sweep itself (first_step).
static int run_stage(gks2d_ctx* c, int stage, double dt, const int32_t* bc, const Dev& upd, bool primvar) {
  const int kinds[4] = {bc[0], bc[1], bc[2], bc[3]};
  RC(timed(c, PH_BC, [&] { return c->lt->bc_all(c->d, kinds, c->s); }));
  if (primvar) RC(timed(c, PH_PRIMVAR, [&] { return c->lt->primvar(c->d, nullptr, stage, 4, c->s); }));
  RC(gks2d_reconstruct(c));
  RC(gks2d_flux(c, stage, dt));
  RC(timed(c, PH_UPDATE, [&] { return c->lt->update_stat(upd, 1 - stage, c->s); }));
  if (stage == 0) {
    HIPCHK(hipMemcpyAsync(c->h_stat, c->d.stat, 8 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipEventRecord(c->ev, c->s));
  }
  if (c->p.use_df && !c->p.df_frozen) RC(gks2d_update_alpha(c));
  return GKS_OK;
}
static int run_step(gks2d_ctx* c, double dt, const int32_t* bc, bool first_step) {
  double* const A = c->d.U;
  double* const B = c->d.Uold;
  Dev upd = c->d;
  upd.U = B;
  upd.Uold = A;
  RC(run_stage(c, 0, dt, bc, upd, first_step));
  c->d.U = B;
  c->d.Uold = A;
  upd = c->d;
  upd.U = A;
  const int rc = run_stage(c, 1, dt, bc, upd, false);
  c->d.U = A;
  c->d.Uold = B;
  return rc;
}

int gks2d_step(gks2d_ctx* c, double dt, const int32_t bc[4]) {
  if (!c || !bc) return fail(GKS_ERR_ARG, "null");
  for (int s = 0; s < 4; s++)
    if (bc[s] < 0 || bc[s] > 2) return fail(GKS_ERR_ARG, "bc kind");
  RC(gks2d_copy_old(c));
  RC(stage_launch(c, 0, dt, bc));
  RC(stage_launch(c, 1, dt, bc));
  c->last_slot = 1;
  c->last_flux_stage = 1;
  return GKS_OK;
}

int gks2d_run(gks2d_ctx* c, const int32_t bc[4], double tstop, int32_t max_steps, double* t_inout,
              int32_t* step_inout) {
  if (!c || !bc || !t_inout || !step_inout) return fail(GKS_ERR_ARG, "null");
  if (c->last_slot < 0) return fail(GKS_ERR_STATE, "no primitive state yet: call gks2d_set_cfl_primvar first");
  for (int s = 0; s < 4; s++)
    if (bc[s] < 0 || bc[s] > 2) return fail(GKS_ERR_ARG, "bc kind");
  double t = *t_inout;
  int32_t step = *step_inout;
  // speed for the first dt: whatever the last primvar produced
  HIPCHK(hipMemcpyAsync(c->h_stat, c->d.stat, 8 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  double tmax = c->h_stat[2 * c->last_slot];
  int n = 0, launched = 0;
  // the last step's stage-1 ghost fill, into the state (run_step) on every
  // way out of the loop below once a step has been launched
  struct GhostsBack {
    gks2d_ctx* c;
    int* launched;
    ~GhostsBack() {
      if (*launched > 0) (void)c->lt->copy_ghosts(c->d, c->d.Uold, c->s);
    }
  } ghosts_back{c, &launched};
  while (t < tstop && n < max_steps) {
    if (c->h_stat[4] != 0.0) break;
    const double dt = dt_from_speed(c, tmax, t, tstop);
    RC(run_step(c, dt, bc, n == 0));
    launched++;
    c->last_slot = 1;
    c->last_flux_stage = 1;
    HIPCHK(hipEventSynchronize(c->ev));
    if (c->h_stat[4] != 0.0) break;
    tmax = c->h_stat[2];
    t = t + dt;
    step++;
    n++;
  }
  *t_inout = t;
  *step_inout = step;
  if (c->h_stat[4] != 0.0) {
    HIPCHK(hipStreamSynchronize(c->s));
    const double zero = 0.0;
    HIPCHK(hipMemcpyAsync(c->d.stat + 4, &zero, sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->h_stat[4] = 0.0;
    return fail(GKS_ERR_BLOWUP, "density or pressure negative or NaN");
  }
  return GKS_OK;
}

int gks2d_keep_face_fluxes(gks2d_ctx* c, int on) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  Dev& d = c->d;
  if (on && !d.flx[0]) {
    const int gp = d.gp;
    int rc;
    if ((rc = dev_alloc(c, &d.flx[0], 8 * gp * d.nfx)) || (rc = dev_alloc(c, &d.flx[1], 8 * gp * d.nfx)) ||
        (rc = dev_alloc(c, &d.fly[0], 8 * gp * d.nfy)) || (rc = dev_alloc(c, &d.fly[1], 8 * gp * d.nfy)))
      return rc;
  }
  d.keep_face_fluxes = on ? 1 : 0;
  c->kept_stage_mask = 0;
  return GKS_OK;
}

int gks2d_get_face_fluxes(gks2d_ctx* c, double* xf, double* yf) {
  if (!c || !xf || !yf) return fail(GKS_ERR_ARG, "null");
  if (c->last_flux_stage < 0) return fail(GKS_ERR_STATE, "no flux computed yet");
  if (!(c->kept_stage_mask & (1 << c->last_flux_stage)))
    return fail(GKS_ERR_STATE, "face fluxes of the last stage were not kept: call gks2d_keep_face_fluxes(ctx, 1) first");
  const Dev& d = c->d;
  const int st = c->last_flux_stage, gp = d.gp;
  std::vector<double> h(8 * gp * d.nfx);
  HIPCHK(hipMemcpyAsync(h.data(), d.flx[st], h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (long long f = 0; f < d.nfx; f++)
    for (int q = 0; q < gp; q++)
      for (int k = 0; k < 8; k++) xf[(f * gp + q) * 8 + k] = h[(q * 8 + k) * d.nfx + f];
  h.assign(8 * gp * d.nfy, 0.0);
  HIPCHK(hipMemcpyAsync(h.data(), d.fly[st], h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (long long f = 0; f < d.nfy; f++)
    for (int q = 0; q < gp; q++)
      for (int k = 0; k < 8; k++) yf[(f * gp + q) * 8 + k] = h[(q * 8 + k) * d.nfy + f];
  return GKS_OK;
}

static int flux_points_impl(const gks2d_params* p, int64_t n, const double* left, const double* right, double dt,
                            double* center, double* f, double* derf, int smooth) {
  if (!p || !left || !right || !f || !derf) return fail(GKS_ERR_ARG, "null");
  if (n < 0) return fail(GKS_ERR_ARG, "n < 0");
  if (n == 0) return GKS_OK;
  const LaunchTable* lt = p->fast_math ? &fast::table : &ieee::table;
  double *dl = nullptr, *dr = nullptr, *dc = nullptr, *dfo = nullptr, *dd = nullptr;
  hipError_t e = hipMalloc(&dl, n * 12 * 8);
  if (e == hipSuccess) e = hipMalloc(&dr, n * 12 * 8);
  if (e == hipSuccess) e = hipMalloc(&dc, n * 12 * 8);
  if (e == hipSuccess) e = hipMalloc(&dfo, n * 4 * 8);
  if (e == hipSuccess) e = hipMalloc(&dd, n * 4 * 8);
  if (e == hipSuccess) e = hipMemcpy(dl, left, n * 12 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dr, right, n * 12 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = lt->flux_points(gas_of(*p), n, dl, dr, dt, dc, dfo, dd, smooth, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && center && !smooth) e = hipMemcpy(center, dc, n * 12 * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(f, dfo, n * 4 * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(derf, dd, n * 4 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(dl);
  (void)hipFree(dr);
  (void)hipFree(dc);
  (void)hipFree(dfo);
  (void)hipFree(dd);
  if (e != hipSuccess) return hip_fail(e, "gks2d_flux_points");
  return GKS_OK;
}
int gks2d_flux_points(const gks2d_params* p, int64_t n, const double* left, const double* right, double dt,
                      double* center, double* f, double* derf) {
  return flux_points_impl(p, n, left, right, dt, center, f, derf, 0);
}
int gks2d_flux_points_smooth(const gks2d_params* p, int64_t n, const double* left, const double* right, double dt,
                             double* f, double* derf) {
  return flux_points_impl(p, n, left, right, dt, nullptr, f, derf, 1);
}
int gks2d_linear_weights_points(const gks2d_params* p, int64_t n, const double* stencils, double h, double* out) {
  if (!p || !stencils || !out) return fail(GKS_ERR_ARG, "null");
  if (n < 0 || !(h > 0.0)) return fail(GKS_ERR_ARG, "n < 0 or h <= 0");
  if (n == 0) return GKS_OK;
  const LaunchTable* lt = p->fast_math ? &fast::table : &ieee::table;
  double *dw = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&dw, n * 5 * 8);
  if (e == hipSuccess) e = hipMalloc(&dout, n * 8 * 8);
  if (e == hipSuccess) e = hipMemcpy(dw, stencils, n * 5 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = lt->linear_points(n, dw, h, dout, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, dout, n * 8 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(dw);
  (void)hipFree(dout);
  if (e != hipSuccess) return hip_fail(e, "gks2d_linear_weights_points");
  return GKS_OK;
}

// ------------------------------------------------- learned classifier hooks
// 2 kernels (normal, tangential) x 1024 slots x (linear, total) counters,
// then the listed flux faces of x and y sweeps
constexpr size_t kLstatPairs = 2 * 1024 * 2;
constexpr size_t kLstatWords = kLstatPairs + 2;
int gks2d_classifier_samples(gks2d_ctx* c, int32_t dir, int32_t stride, double label_tol, float* features,
                             float* labels, int64_t cap, int64_t* n_out) {
  if (!c || !features || !labels || !n_out) return fail(GKS_ERR_ARG, "null");
  if (dir < 0 || dir > 1 || stride < 1) return fail(GKS_ERR_ARG, "dir / stride");
  if (!(label_tol > 0.0)) return fail(GKS_ERR_ARG, "label_tol must be > 0");
  if (c->p.weno_order != 5) return fail(GKS_ERR_UNSUPPORTED, "the classifier is for WENO5-AO");
  const Dev& d = c->d;
  const long long total = dir == 0 ? (long long)(d.NX + 2) * d.ny : (long long)d.nx * (d.NY + 2);
  const long long ns = (total + stride - 1) / stride;
  if (ns > cap) return fail(GKS_ERR_ARG, "cap smaller than the number of samples");
  float *dX = nullptr, *dY = nullptr;
  hipError_t e = hipMalloc(&dX, ns * 16 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dY, ns * sizeof(float));
  if (e == hipSuccess) e = c->lt->label(d, dir, stride, label_tol, dX, dY, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e == hipSuccess) e = hipMemcpy(features, dX, ns * 16 * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(labels, dY, ns * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dX);
  (void)hipFree(dY);
  if (e != hipSuccess) return hip_fail(e, "gks2d_classifier_samples");
  *n_out = ns;
  return GKS_OK;
}

// learned-path work lists (each in kListShards shards): the shard counts
// of the flux lists x, y and of the normal lists x, y (one 128-byte line
// each), the four lists' prefix records (kListRecord words each), then the
// flux faces x, y and the normal-sweep cells x, y.  A shard takes the
// appends of a run of at most ceil(tiles / kListShards) tiles of the sweep
// that fills the list (flux tiles: 64 or 128 faces).
// entries one shard can receive from a sweep of `tiles` tiles of per_tile
// items (a shard takes a run of at most ceil(tiles / kListShards) tiles)
static unsigned shard_cap_tiles(long long tiles, int per_tile) {
  return (unsigned)((tiles + kListShards - 1) / kListShards * per_tile);
}
static unsigned shard_cap(long long items, int per_tile) {
  return shard_cap_tiles((items + per_tile - 1) / per_tile, per_tile);
}
static void list_caps(Dev& d) {
  const long long nf[2] = {(long long)(d.NX + 1) * d.NY, (long long)d.NX * (d.NY + 1)};
  const long long ncl[2] = {(long long)(d.NX + 2) * d.ny, (long long)d.nx * (d.NY + 2)};
  // the normal lists: filled by one sweep per direction, or by both
  // directions' classifier launch in tiles of kSweepBlock cells of the frame
  for (int dir = 0; dir < 2; dir++)
    d.ncap[dir] = std::max(shard_cap(ncl[dir], kSweepBlock), shard_cap_tiles(cls_tiles(d), kSweepBlock));
  // the flux lists are filled by sweeps of 64 or 128 faces per tile, in
  // 2D tiles (cheap_tiles); runs along the rows for the single learned
  // launch
  for (int dir = 0; dir < 2; dir++) {
    d.tcap[dir] = 0;
    for (int tf : {64, 128})
      d.tcap[dir] = std::max({d.tcap[dir], shard_cap(nf[dir], tf), shard_cap_tiles(cheap_tiles(d, dir, tf), tf)});
  }
}
// the directions whose linear line values the smooth-face flux launch forms
// itself: it runs when the smooth-flow flux is on (and the tangential
// reconstruction is fused into the flux launch)
static int fused_line_dirs(const Dev& d) {
  return d.smooth_flux && d.fuse_tanflux && d.order == 5 ? kFusedLineDirs : 0;
}
static size_t list_words(const Dev& d) {
  return (size_t)4 * kListShards * kCountStride + 4 * kListRecord +
         (size_t)kListShards * ((size_t)d.tcap[0] + d.tcap[1] + d.ncap[0] + d.ncap[1]);
}
int gks2d_set_classifier(gks2d_ctx* c, const float* params, float logit_threshold) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  HIPCHK(hipStreamSynchronize(c->s));
  if (!params) {
    c->d.mlp = nullptr;
    return GKS_OK;
  }
  if (c->p.weno_order != 5) return fail(GKS_ERR_UNSUPPORTED, "the classifier is for WENO5-AO");
  const size_t nc = (size_t)c->d.nc;
  list_caps(c->d);
  if (!c->d_mlp) {
    HIPCHK(hipMalloc(&c->d_mlp, GKS_MLP_NPARAM * sizeof(float)));
    HIPCHK(hipMalloc(&c->d_mlpw, mlp::kPackHalves * sizeof(unsigned short)));
    HIPCHK(hipMalloc(&c->d_lstat, kLstatWords * sizeof(unsigned long long)));
    HIPCHK(hipMalloc(&c->d_lin, 2 * nc));
    HIPCHK(hipMalloc(&c->d_tlist, list_words(c->d) * sizeof(unsigned)));
  }
  std::vector<uint16_t> packed(mlp::kPackHalves);
  mlp::pack_bf16(params, packed.data());
  HIPCHK(hipMemcpy(c->d_mlp, params, GKS_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_mlpw, packed.data(), packed.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->d_lstat, 0, kLstatWords * sizeof(unsigned long long)));
  HIPCHK(hipMemset(c->d_lin, 0, 2 * nc));
  // the shard counts start cleared; each list's prefix kernel clears them
  // again for the next sweep
  HIPCHK(hipMemset(c->d_tlist, 0, (size_t)4 * kListShards * kCountStride * sizeof(unsigned)));
  c->d.mlp = c->d_mlp;
  c->d.mlpw = c->d_mlpw;
  c->d.smooth_flux = 1;
  c->d.lin[0] = c->d_lin;
  c->d.lin[1] = c->d_lin + nc;
  c->d.tcount = c->d_tlist;
  c->d.ncount = c->d.tcount + 2 * kListShards * kCountStride;
  c->d.lpre = c->d.tcount + 4 * kListShards * kCountStride;
  c->d.tlist[0] = c->d.lpre + 4 * kListRecord;
  c->d.tlist[1] = c->d.tlist[0] + (size_t)kListShards * c->d.tcap[0];
  c->d.nlist[0] = c->d.tlist[1] + (size_t)kListShards * c->d.tcap[1];
  c->d.nlist[1] = c->d.nlist[0] + (size_t)kListShards * c->d.ncap[0];
  c->d.mlp_thr = logit_threshold;
  c->d.lstat = c->d_lstat;
  c->d.fused_lines = fused_line_dirs(c->d);
  return GKS_OK;
}

int gks2d_classifier_options(gks2d_ctx* c, int32_t smooth_flux) {
  if (!c) return fail(GKS_ERR_ARG, "null");
  if (!c->d.mlp) return fail(GKS_ERR_STATE, "no classifier set");
  HIPCHK(hipStreamSynchronize(c->s));
  c->d.smooth_flux = smooth_flux ? 1 : 0;
  c->d.fused_lines = fused_line_dirs(c->d);
  return GKS_OK;
}

int gks2d_classifier_stats(gks2d_ctx* c, int64_t* linear, int64_t* total) {
  if (!c || !linear || !total) return fail(GKS_ERR_ARG, "null");
  if (!c->d_lstat) return fail(GKS_ERR_STATE, "no classifier set");
  std::vector<unsigned long long> h(kLstatWords);
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(h.data(), c->d_lstat, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  unsigned long long lin = 0, tot = 0;
  for (size_t k = 0; k < kLstatPairs; k += 2) {
    lin += h[k];
    tot += h[k + 1];
  }
  *linear = (int64_t)lin;
  *total = (int64_t)tot;
  return GKS_OK;
}

int gks2d_classifier_counts(gks2d_ctx* c, int64_t counts[6]) {
  if (!c || !counts) return fail(GKS_ERR_ARG, "null");
  if (!c->d_lstat) return fail(GKS_ERR_STATE, "no classifier set");
  std::vector<unsigned long long> h(kLstatWords);
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(h.data(), c->d_lstat, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int k = 0; k < 6; k++) counts[k] = 0;
  for (size_t k = 0; k < kLstatPairs; k += 2) {
    const int part = k < kLstatPairs / 2 ? 0 : 2;  // normal, tangential
    counts[part] += (int64_t)h[k];
    counts[part + 1] += (int64_t)h[k + 1];
  }
  counts[4] = (int64_t)h[kLstatPairs];
  counts[5] = (int64_t)h[kLstatPairs + 1];
  return GKS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches
#include "../../include/gks2d_batch.h"

struct gks2d_batch {
  std::vector<gks2d_ctx*> ctx;
  int np = 0;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Dev* d_dev = nullptr;
  BState* d_st = nullptr;
  unsigned* d_start = nullptr;
  int* d_pid = nullptr;
  unsigned* d_lstart = nullptr;  // list tile prefixes (BatchArgs::lstart)
  BState* h_st = nullptr;  // pinned
  BatchShape shape{128, 0};
  int largest_first = 1;
  std::vector<unsigned> tiles = std::vector<unsigned>(KB_COUNT, 0);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_key = -1;  // shape the graph was captured for
  // problems the tile tables (and so the captured step) cover: every
  // problem when a run starts; a problem that finishes is dropped at the
  // next host check, so that the launches stop carrying its tiles
  std::vector<char> in_tables;
  int np_tab = 0;
  bool tables_dirty = true;
};

namespace {

unsigned ceil_div(long long n, long long k) { return (unsigned)((n + k - 1) / k); }

void drop_graph(gks2d_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key = -1;
}

// tile prefix tables of every work class, in launch order
int batch_tables(gks2d_batch* b, int learned) {
  std::vector<int> pid;
  for (int p = 0; p < b->np; p++)
    if (b->in_tables[p]) pid.push_back(p);
  const int np = (int)pid.size();
  b->np_tab = np;
  if (b->largest_first)
    std::stable_sort(pid.begin(), pid.end(), [&](int a, int c) { return b->ctx[a]->d.nc > b->ctx[c]->d.nc; });
  std::vector<unsigned> start((size_t)KB_COUNT * (np + 1));
  const int tf = b->shape.tanflux_faces;
  for (int k = 0; k < KB_COUNT; k++) {
    unsigned acc = 0;
    for (int q = 0; q < np; q++) {
      start[(size_t)k * (np + 1) + q] = acc;
      const Dev& d = b->ctx[pid[q]]->d;
      long long n = 0;
      switch (k) {
        case KB_BC_LR: n = ceil_div((long long)d.G * d.ny, kSweepBlock); break;
        case KB_BC_DU: n = ceil_div((long long)d.G * d.nx, kSweepBlock); break;
        case KB_UPDATE: n = ceil_div((long long)d.NX * d.NY, kSweepBlock); break;
        // the learned path's classifier launch takes both directions
        case KB_NORMAL_X: n = learned ? cls_tiles(d) : ceil_div((long long)(d.NX + 2) * d.ny, kSweepBlock); break;
        case KB_NORMAL_Y: n = ceil_div((long long)d.nx * (d.NY + 2), kSweepBlock); break;
        // the learned path's smooth-face launch takes 2D tiles
        case KB_FLUX_X: n = learned ? cheap_tiles(d, 0, tf) : ceil_div((long long)(d.NX + 1) * d.NY, tf); break;
        case KB_FLUX_Y: n = learned ? cheap_tiles(d, 1, tf) : ceil_div((long long)d.NX * (d.NY + 1), tf); break;
        case KB_ALPHA: n = ceil_div((long long)d.NX * std::min(d.NX, d.NY), kSweepBlock); break;
      }
      if (acc + (unsigned long long)n > 0x7fffffffULL) return fail(GKS_ERR_ARG, "batch too large for one launch");
      acc += (unsigned)n;
    }
    start[(size_t)k * (np + 1) + np] = acc;
    b->tiles[k] = acc;
  }
  HIPCHK(hipMemcpy(b->d_start, start.data(), start.size() * sizeof(unsigned), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_pid, pid.data(), pid.size() * sizeof(int), hipMemcpyHostToDevice));
  return GKS_OK;
}

// device copies of the problems' descriptors (they change when a
// classifier is set) and a consistency check of the batch
int batch_upload_devs(gks2d_batch* b, int* learned) {
  std::vector<Dev> h(b->np);
  for (int p = 0; p < b->np; p++) {
    gks2d_ctx* c = b->ctx[p];
    HIPCHK(hipStreamSynchronize(c->s));
    h[p] = c->d;
    // the batch's learned sweep always runs the smooth-face flux launch
    if (h[p].mlp && h[p].order == 5) h[p].fused_lines = kFusedLineDirs;
    if (c->last_slot < 0) return fail(GKS_ERR_STATE, "batch problem without a primitive state: set its initial condition first");
  }
  const int l0 = h[0].mlp != nullptr;
  for (int p = 1; p < b->np; p++)
    if ((h[p].mlp != nullptr) != l0) return fail(GKS_ERR_STATE, "the classifier must be set on all problems or on none");
  *learned = l0;
  HIPCHK(hipMemcpy(b->d_dev, h.data(), h.size() * sizeof(Dev), hipMemcpyHostToDevice));
  return GKS_OK;
}

int batch_init_state(gks2d_batch* b, const int32_t* bc, const double* tstop, int32_t max_steps, const double* t,
                     const int32_t* steps) {
  for (int p = 0; p < b->np; p++) {
    BState& st = b->h_st[p];
    std::memset(&st, 0, sizeof(st));
    st.t = t[p];
    st.tstop = tstop[p];
    st.cfl = b->ctx[p]->p.cfl;
    st.slot = b->ctx[p]->last_slot;
    st.steps = steps[p];
    st.max_steps = (int)std::min<long long>((long long)steps[p] + max_steps, 0x7fffffff);
    for (int k = 0; k < 4; k++) {
      if (bc[4 * p + k] < 0 || bc[4 * p + k] > 2) return fail(GKS_ERR_ARG, "bc kind");
      st.bc[k] = bc[4 * p + k];
    }
  }
  HIPCHK(hipMemcpyAsync(b->d_st, b->h_st, b->np * sizeof(BState), hipMemcpyHostToDevice, b->s));
  // the initial states' blow-up test (in the batched step the updates test
  // the states they write), after the boundary fill its ghost cells need
  // (the first stage fills them the same way): raised into stat[5], which
  // the first step's end takes, as that step's stage-0 primvar sweep would
  // have; no speed slot is written (slot -1: the first dt reads the speed
  // the problem's last primvar formed)
  for (int p = 0; p < b->np; p++) {
    const Dev& d = b->ctx[p]->d;
    const int kinds[4] = {bc[4 * p], bc[4 * p + 1], bc[4 * p + 2], bc[4 * p + 3]};
    HIPCHK(hipMemsetAsync(d.stat + 5, 0, 2 * sizeof(double), b->s));
    HIPCHK(b->ctx[p]->lt->bc_all(d, kinds, b->s));
    HIPCHK(b->ctx[p]->lt->primvar(d, nullptr, -1, 5, b->s));
  }
  return GKS_OK;
}

// the captured step for the current shape
int batch_graph(gks2d_batch* b, int learned) {
  const int key = b->shape.tanflux_faces * 4 + b->largest_first * 2 + learned;
  if (b->exec && b->graph_key == key && !b->tables_dirty) return GKS_OK;
  drop_graph(b);
  RC(batch_tables(b, learned));
  b->tables_dirty = false;
  b->shape.learned = learned;
  const BatchArgs a{b->d_dev, b->d_st, b->d_start, b->d_pid, b->np_tab, b->d_lstart};
  const LaunchTable* lt = b->ctx[0]->lt;
  HIPCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
  hipError_t e = lt->batch_step(a, b->shape, b->tiles.data(), b->s);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->s, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(e, "batch_step capture");
  }
  if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph_key = key;
  return GKS_OK;
}

// the tile tables cover every problem again (start of a run, autotuning)
void batch_cover_all(gks2d_batch* b) {
  for (int p = 0; p < b->np; p++)
    if (!b->in_tables[p]) {
      b->in_tables[p] = 1;
      b->tables_dirty = true;
    }
}

bool batch_all_done(const gks2d_batch* b) {
  for (int p = 0; p < b->np; p++)
    if (!b->h_st[p].done) return false;
  return true;
}

}  // namespace

extern "C" {

int gks2d_batch_destroy(gks2d_batch* b) {
  if (!b) return GKS_OK;
  if (b->s) (void)hipStreamSynchronize(b->s);
  drop_graph(b);
  for (gks2d_ctx* c : b->ctx) gks2d_destroy(c);
  if (b->d_dev) (void)hipFree(b->d_dev);
  if (b->d_st) (void)hipFree(b->d_st);
  if (b->d_start) (void)hipFree(b->d_start);
  if (b->d_pid) (void)hipFree(b->d_pid);
  if (b->d_lstart) (void)hipFree(b->d_lstart);
  if (b->h_st) (void)hipHostFree(b->h_st);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
  return GKS_OK;
}

int gks2d_batch_create(const gks2d_params* params, int32_t n, gks2d_batch** out) {
  if (!params || !out) return fail(GKS_ERR_ARG, "null");
  *out = nullptr;
  if (n < 1 || n > 1024) return fail(GKS_ERR_ARG, "a batch holds 1 to 1024 problems");
  for (int p = 0; p < n; p++) {
    if (params[p].weno_order != 5) return fail(GKS_ERR_UNSUPPORTED, "batches run WENO5-AO (2 Gauss points)");
    if (params[p].fast_math != params[0].fast_math) return fail(GKS_ERR_ARG, "fast_math must agree in a batch");
  }
  gks2d_batch* b = new gks2d_batch();
  b->np = n;
  b->in_tables.assign(n, 1);
  for (int p = 0; p < n; p++) {
    gks2d_ctx* c = nullptr;
    int rc = gks2d_create(&params[p], &c);
    if (rc) {
      gks2d_batch_destroy(b);
      return rc;
    }
    if (!c->d.fuse_tanflux) {
      gks2d_destroy(c);
      gks2d_batch_destroy(b);
      return fail(GKS_ERR_UNSUPPORTED, "batches need the fused tangential + flux launch (GKS_TANFLUX=0 is set)");
    }
    b->ctx.push_back(c);
  }
  hipError_t e = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->e0);
  if (e == hipSuccess) e = hipEventCreate(&b->e1);
  if (e == hipSuccess) e = hipMalloc(&b->d_dev, n * sizeof(Dev));
  if (e == hipSuccess) e = hipMalloc(&b->d_st, n * sizeof(BState));
  if (e == hipSuccess) e = hipMalloc(&b->d_start, (size_t)KB_COUNT * (n + 1) * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc(&b->d_pid, n * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&b->d_lstart, (size_t)KL_COUNT * (n + 1) * sizeof(unsigned));
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_st, n * sizeof(BState));
  if (e != hipSuccess) {
    gks2d_batch_destroy(b);
    return hip_fail(e, "gks2d_batch_create");
  }
  *out = b;
  return GKS_OK;
}

int gks2d_batch_problem(gks2d_batch* b, int32_t p, gks2d_ctx** out) {
  if (!b || !out) return fail(GKS_ERR_ARG, "null");
  if (p < 0 || p >= b->np) return fail(GKS_ERR_ARG, "problem index");
  *out = b->ctx[p];
  return GKS_OK;
}

int gks2d_batch_set_shape(gks2d_batch* b, int32_t tanflux_faces, int32_t largest_first) {
  if (!b) return fail(GKS_ERR_ARG, "null");
  if (tanflux_faces != 64 && tanflux_faces != 128) return fail(GKS_ERR_ARG, "tanflux_faces: 64 or 128");
  b->shape.tanflux_faces = tanflux_faces;
  b->largest_first = largest_first ? 1 : 0;
  return GKS_OK;
}

int gks2d_batch_run(gks2d_batch* b, const int32_t* bc, const double* tstop, int32_t max_steps,
                    int32_t check_every, double* t, int32_t* steps) {
  if (!b || !bc || !tstop || !t || !steps) return fail(GKS_ERR_ARG, "null");
  if (max_steps < 0) return fail(GKS_ERR_ARG, "max_steps");
  if (check_every < 1) check_every = 1;
  int learned = 0;
  RC(batch_upload_devs(b, &learned));
  RC(batch_init_state(b, bc, tstop, max_steps, t, steps));
  const std::vector<int32_t> steps0(steps, steps + b->np);
  batch_cover_all(b);
  RC(batch_graph(b, learned));
  // every replay advances each unfinished problem by one step (or ends it),
  // so max_steps + 1 replays end every problem
  long long left = (long long)max_steps + 1;
  while (left > 0) {
    const long long r = std::min<long long>(check_every, left);
    for (long long k = 0; k < r; k++) HIPCHK(hipGraphLaunch(b->exec, b->s));
    left -= r;
    HIPCHK(hipMemcpyAsync(b->h_st, b->d_st, b->np * sizeof(BState), hipMemcpyDeviceToHost, b->s));
    HIPCHK(hipStreamSynchronize(b->s));
    if (batch_all_done(b)) break;
    // drop the problems that finished from the launches (new tables, one
    // more capture; a finished problem's tiles would otherwise still be
    // dispatched every replay only to return)
    bool shrink = false;
    for (int p = 0; p < b->np; p++)
      if (b->in_tables[p] && b->h_st[p].done) {
        b->in_tables[p] = 0;
        shrink = true;
      }
    if (shrink) {
      b->tables_dirty = true;
      RC(batch_graph(b, learned));
    }
  }
  std::string bad;
  for (int p = 0; p < b->np; p++) {
    const BState& st = b->h_st[p];
    t[p] = st.t;
    steps[p] = st.steps;
    gks2d_ctx* c = b->ctx[p];
    // a problem that ran a step: its last stage-1 ghost fill, into the state
    // (batch_stage: stage 1 filled the second buffer's ghost frame)
    if (st.steps > steps0[p] || st.blowup) HIPCHK(c->lt->copy_ghosts(c->d, c->d.Uold, b->s));
    c->last_slot = st.slot;
    if (st.steps > 0) c->last_flux_stage = 1;
    if (st.blowup) {
      const double zero = 0.0;
      HIPCHK(hipMemcpy(c->d.stat + 4, &zero, sizeof(double), hipMemcpyHostToDevice));
      bad += (bad.empty() ? "" : ", ") + std::to_string(p);
    }
  }
  // the ghost copies above ran on the batch's stream
  HIPCHK(hipStreamSynchronize(b->s));
  if (!bad.empty()) return fail(GKS_ERR_BLOWUP, ("density or pressure negative or NaN in problem(s) " + bad).c_str());
  return GKS_OK;
}

int gks2d_batch_autotune(gks2d_batch* b, const int32_t* bc, int32_t steps, int32_t* tanflux_faces,
                         int32_t* largest_first, double* ms_per_step) {
  if (!b || !bc) return fail(GKS_ERR_ARG, "null");
  if (steps < 1) return fail(GKS_ERR_ARG, "steps");
  int learned = 0;
  RC(batch_upload_devs(b, &learned));
  batch_cover_all(b);
  // snapshot of what a step changes and a later step reads
  std::vector<double*> snap(b->np, nullptr);
  std::vector<unsigned long long*> lsnap(b->np, nullptr);
  auto release = [&]() {
    for (double* q : snap)
      if (q) (void)hipFree(q);
    for (unsigned long long* q : lsnap)
      if (q) (void)hipFree(q);
  };
  auto copy_state = [&](bool save) -> hipError_t {
    for (int p = 0; p < b->np; p++) {
      const Dev& d = b->ctx[p]->d;
      double* q = snap[p];
      const size_t nc = (size_t)d.nc;
      struct {
        double* dev;
        size_t n;
      } parts[4] = {{d.U, 4 * nc}, {d.ax, nc}, {d.ay, nc}, {d.stat, 8}};
      size_t off = 0;
      for (auto& part : parts) {
        hipError_t e = save ? hipMemcpyAsync(q + off, part.dev, part.n * 8, hipMemcpyDeviceToDevice, b->s)
                            : hipMemcpyAsync(part.dev, q + off, part.n * 8, hipMemcpyDeviceToDevice, b->s);
        if (e != hipSuccess) return e;
        off += part.n;
      }
      if (lsnap[p]) {
        hipError_t e = save ? hipMemcpyAsync(lsnap[p], d.lstat, kLstatWords * 8, hipMemcpyDeviceToDevice, b->s)
                            : hipMemcpyAsync(d.lstat, lsnap[p], kLstatWords * 8, hipMemcpyDeviceToDevice, b->s);
        if (e != hipSuccess) return e;
      }
    }
    return hipSuccess;
  };
  for (int p = 0; p < b->np; p++) {
    const Dev& d = b->ctx[p]->d;
    hipError_t e = hipMalloc(&snap[p], (6 * (size_t)d.nc + 8) * sizeof(double));
    if (e == hipSuccess && d.lstat) e = hipMalloc(&lsnap[p], kLstatWords * sizeof(unsigned long long));
    if (e != hipSuccess) {
      release();
      return hip_fail(e, "gks2d_batch_autotune snapshot");
    }
  }
  std::vector<double> t0(b->np, 0.0), big(b->np, 1e300);
  std::vector<int32_t> s0(b->np, 0);
  hipError_t e = copy_state(true);
  const BatchShape keep = b->shape;
  const int keep_order = b->largest_first;
  double best = 1e300;
  int best_tf = keep.tanflux_faces, best_order = keep_order;
  int rc = GKS_OK;
  const int tfs[2] = {64, 128};
  for (int ti = 0; ti < 2 && e == hipSuccess && rc == GKS_OK; ti++) {
    for (int order = 0; order < 2 && e == hipSuccess && rc == GKS_OK; order++) {
      b->shape.tanflux_faces = tfs[ti];
      b->largest_first = order;
      if ((rc = batch_init_state(b, bc, big.data(), steps + 1, t0.data(), s0.data()))) break;
      if ((rc = batch_graph(b, learned))) break;
      e = hipGraphLaunch(b->exec, b->s);  // warm-up step
      if (e == hipSuccess) e = hipEventRecord(b->e0, b->s);
      for (int k = 0; k < steps && e == hipSuccess; k++) e = hipGraphLaunch(b->exec, b->s);
      if (e == hipSuccess) e = hipEventRecord(b->e1, b->s);
      if (e == hipSuccess) e = hipEventSynchronize(b->e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->e0, b->e1);
      if (e == hipSuccess && ms / steps < best) {
        best = ms / steps;
        best_tf = tfs[ti];
        best_order = order;
      }
      if (e == hipSuccess) e = copy_state(false);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(b->s);
  release();
  if (rc) {
    b->shape = keep;
    b->largest_first = keep_order;
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "gks2d_batch_autotune");
  b->shape.tanflux_faces = best_tf;
  b->largest_first = best_order;
  if (tanflux_faces) *tanflux_faces = best_tf;
  if (largest_first) *largest_first = best_order;
  if (ms_per_step) *ms_per_step = best;
  return GKS_OK;
}

}  // extern "C"
