// lj_api.cpp -- the C-ABI of libljhip.so: compatibility set (the reference's
// MDSystem.cu entry points), context API and device-pointer API.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lj_ctx.h"
#include "lj_fp64.h"

using namespace lj;

namespace {
thread_local std::string g_err;
}  // namespace

namespace lj {
int fail(int code, const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
const std::string& last_error() { return g_err; }
thread_local LaunchStamp tl_force_stamp;
}  // namespace lj

namespace lj {

int ctx_setdev(lj_ctx* c)
{
  if (!c) return fail(LJ_ERR_ARG, "null context");
  LJ_HIP(hipSetDevice(c->cfg.device));
  return LJ_OK;
}

void drop_graph(lj_ctx* c)
{
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->gexec_run) (void)hipGraphExecDestroy(c->gexec_run);
  if (c->graph_run) (void)hipGraphDestroy(c->graph_run);
  if (c->gexec_long) (void)hipGraphExecDestroy(c->gexec_long);
  if (c->graph_long) (void)hipGraphDestroy(c->graph_long);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->gexec_run = nullptr;
  c->graph_run = nullptr;
  c->gexec_long = nullptr;
  c->graph_long = nullptr;
}

// rows of one of the two block-row buffers
size_t blk_rows(const lj_ctx* c) { return (size_t)std::max(c->nblk, step_rows(c->cfg.N)); }

// b: which of the two buffers of per-workgroup totals (the step's force pass)
// and of block rows (its integrator kernels) the step fills
StepArgs step_args(lj_ctx* c, double dt, int b = 0)
{
  StepArgs a{};
  a.pos = c->pos;
  a.vel = c->vel;
  a.force = c->force;
  a.tF = c->tF;
  a.q = c->q;
  a.fpart = c->fpart;
  a.apart = c->apart;
  a.blk = c->blk + (is_fp64(c) ? 0 : (size_t)b * blk_rows(c) * kBlkFields);
  a.obs = c->obs;
  a.hist = c->hist;
  a.hist_base = c->hist_base;
  a.hist_cap = c->hist_cap;
  a.N = c->cfg.N;
  a.dt = dt;
  a.L = c->L;
  a.T0 = c->cfg.T0;
  a.rho = c->cfg.rho;
  a.eps = c->cfg.eps;
  a.qscale = c->ps.qscale;
  a.FS = c->ps.FS;
  a.bc = c->cfg.boundary;
  a.canonical = c->cfg.canonical;
  a.S = fplan(c).S;
  a.Npad = fplan(c).Npad;
  a.ew = c->step_ew ? c->step_ew + (size_t)b * 2 * c->plan3.step_wgs() : nullptr;
  a.nitems = c->step_ew ? c->plan3.step_wgs() : 0;
  a.nrow = is_fp64(c) ? c->nblk : step_rows(c->cfg.N);
  return a;
}

void refresh_scalars(lj_ctx* c)
{
  c->ps = make_pair_scalars(c->L, c->cfg.boundary == LJ_BC_PERIODIC, c->cfg.eps, c->cfg.sigma,
                            false);
}

Step64 step64_args(lj_ctx* c, double dt)
{
  Step64 a{};
  a.x = c->xd;
  a.v = c->vd;
  a.f = c->fd;
  a.tf = c->tfd;
  a.pos = c->pos;
  a.vel = c->vel;
  a.force = c->force;
  a.part = c->part64;
  a.apart = c->apart64;
  a.blk = c->blk;
  a.N = c->cfg.N;
  a.S = c->plan.S;
  a.Npad = c->plan.Npad;
  a.bc = c->cfg.boundary;
  a.dt = dt;
  a.L = c->L;
  a.T0 = c->cfg.T0;
  a.FS = 24. * c->cfg.eps / (c->cfg.sigma * c->cfg.sigma);
  return a;
}

double4* state64(lj_ctx* c, int type)
{
  return type == LJ_ARRAY_POS ? c->xd : type == LJ_ARRAY_VEL ? c->vd : c->fd;
}

hipError_t enqueue_force64(lj_ctx* c, hipStream_t st)
{
  return launch_force64(st, c->plan, c->xd, c->cfg.boundary == LJ_BC_PERIODIC, c->L,
                        c->cfg.sigma, c->part64, c->apart64);
}

// the FP64 step: MDSystem.cpp:442-510 in double
int enqueue_step64(lj_ctx* c, hipStream_t st, double dt)
{
  const Step64 a = step64_args(c, dt);
  LJ_HIP(launch_drift64(st, a, c->cfg.canonical));
  LJ_HIP(enqueue_force64(c, st));
  if (c->cfg.canonical == LJ_EVN) {
    LJ_HIP(launch_kick64(st, a));
  } else {
    LJ_HIP(launch_tvn64_a(st, a));
    LJ_HIP(launch_tvn64_b(st, a));
  }
  LJ_HIP(launch_finalize_step(st, step_args(c, dt)));
  // the float mirrors the step kernels wrote are the sampled state
  if (c->smp_fl) LJ_HIP(launch_fl_sample(st, c->fl, c->pos, c->vel, c->cfg.N, nullptr, nullptr, 0));
  return LJ_OK;
}

// force pass into fpart/apart from the context positions (or external d_pos)
int enqueue_force(lj_ctx* c, hipStream_t st, const float4* d_pos)
{
  if (c->ps.periodic) LJ_HIP(launch_prep_q(st, d_pos, c->q, c->cfg.N, c->ps.qscale));
  LJ_HIP(launch_force(st, fplan(c), c->ps, c->q, d_pos, c->fpart, c->apart,
                      c->cfg.precision == LJ_PREC_COMPENSATED, true));
  return LJ_OK;
}

// One step of a run of steps captured together.  The first step starts with
// the drift; each later one finds it done by the previous step's last kernel
// (drift_next).  With the pair-once pass the observables of a step are
// computed by one extra block of the next step's first integrator kernel, and
// only the last step launches the stand-alone finaliser.  Steps alternate
// between the two buffers of per-workgroup totals and of block rows (b = step
// parity), so that the finaliser of step k reads its sums while step k + 1
// writes its own.  Every kernel does the arithmetic of the one-step sequence,
// so a run gives the bits of single steps.
int enqueue_step_part(lj_ctx* c, hipStream_t st, double dt, bool first, bool last, int b)
{
  if (is_fp64(c)) return enqueue_step64(c, st, dt);
  const StepArgs a = step_args(c, dt, b);
  const bool comp = c->cfg.precision == LJ_PREC_COMPENSATED;
  const bool fused_fin = a.nitems > 0;  // the pair-once time-step pass
  if (first) LJ_HIP(launch_drift(st, a));
  // the drift kernel already wrote the fixed-point copy for periodic boxes
  double* ew = const_cast<double*>(a.ew);
  const ForcePlan& p3 = c->plan3;
  if (c->emulate_shards > 1 && p3.n3 == 2 && p3.split == 1 && !p3.comp) {
    // each item writes its own partial entries and totals: the G launches
    // leave the buffers of the single launch
    const int M = p3.nitems_step, G = std::min(c->emulate_shards, M);
    for (int g = 0; g < G; ++g) {
      const int f = (int)((long)g * M / G), e = (int)((long)(g + 1) * M / G);
      LJ_HIP(launch_force_n3_step_items(st, p3, c->ps, c->q, c->pos, c->fpart, ew, f, e - f));
    }
  } else {
    LJ_HIP(launch_force(st, fplan(c), c->ps, c->q, c->pos, c->fpart, c->apart, comp, false, ew));
  }
  if (c->smp_rdf != LJ_SAMPLE_RDF_OFF) {
    // the positions the force pass saw (drifted, before the boundary
    // conditions), as CalculateForces bins them inside Integrate
    if (c->smp_rdf == LJ_SAMPLE_RDF_LAST) LJ_HIP(launch_fill_u64(st, c->rdf_step, c->smp_args.nbins));
    LJ_HIP(launch_rdf(st, c->smp_args));
  }
  const StepArgs prev = step_args(c, dt, b ^ 1);
  const StepArgs* fin = fused_fin && !first ? &prev : nullptr;
  if (c->cfg.canonical == LJ_EVN) {
    LJ_HIP(launch_kick_evn(st, a, !last, fin));
  } else {
    LJ_HIP(launch_tvn_a(st, a, fin));
    LJ_HIP(launch_tvn_b(st, a, !last));
  }
  if (last || !fused_fin) LJ_HIP(launch_finalize_step(st, a));
  if (c->smp_fl) LJ_HIP(launch_fl_sample(st, c->fl, c->pos, c->vel, c->cfg.N, nullptr, nullptr, 0));
  return LJ_OK;
}

int enqueue_step(lj_ctx* c, hipStream_t st, double dt)
{
  return enqueue_step_part(c, st, dt, true, true, 0);
}

int ensure_hist(lj_ctx* c, int n)
{
  if (n <= c->hist_cap) return LJ_OK;
  int cap = std::max(n, 2 * c->hist_cap);
  double* h = nullptr;
  LJ_TRY(dev_alloc(&h, (size_t)5 * cap));
  LJ_HIP(hipStreamSynchronize(c->stream));
  dev_free(c->hist);
  c->hist = h;
  c->hist_cap = cap;
  drop_graph(c);  // the graph holds the old pointer
  return LJ_OK;
}

namespace {
// captures `steps` steps enqueued back to back into one executable graph
int capture_steps(lj_ctx* c, double dt, int steps, hipGraph_t* graph, hipGraphExec_t* exec)
{
  hipStream_t st = c->stream;
  LJ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = LJ_OK;
  // steps that sample end with their own final state (no drift of the next
  // step fused into them)
  const bool alone = c->sampling();
  for (int k = 0; k < steps && rc == LJ_OK; ++k)
    rc = enqueue_step_part(c, st, dt, alone || k == 0, alone || k == steps - 1, k & 1);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(st, &g);
  if (rc != LJ_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  LJ_HIP(ee);
  *graph = g;
  LJ_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return LJ_OK;
}
}  // namespace

// LJ_LONG_GRAPH=0 keeps runs to the 8-step and single-step graphs: under
// rocprofv3's kernel trace the 32-step replays left the device idle about 80 us
// per step (N = 16384), so step traces are taken without them
static bool long_graph_enabled()
{
  const char* e = std::getenv("LJ_LONG_GRAPH");
  return !(e && e[0] == '0');
}

// the step captured once per (dt, ensemble, boundary) and replayed
int ensure_graph(lj_ctx* c, double dt)
{
  if (c->gexec && c->g_dt == dt && c->g_canon == c->cfg.canonical && c->g_bc == c->cfg.boundary)
    return LJ_OK;
  drop_graph(c);
  LJ_HIP(hipStreamSynchronize(c->stream));
  LJ_TRY(capture_steps(c, dt, 1, &c->graph, &c->gexec));
  int rc = capture_steps(c, dt, kGraphSteps, &c->graph_run, &c->gexec_run);
  if (rc == LJ_OK && long_graph_enabled())
    rc = capture_steps(c, dt, kLongGraphSteps, &c->graph_long, &c->gexec_long);
  if (rc != LJ_OK) {
    drop_graph(c);
    return rc;
  }
  c->g_dt = dt;
  c->g_canon = c->cfg.canonical;
  c->g_bc = c->cfg.boundary;
  return LJ_OK;
}

int launch_steps(lj_ctx* c, int nsteps)
{
  int k = 0;
  if (c->gexec_long)
    for (; k + kLongGraphSteps <= nsteps; k += kLongGraphSteps)
      LJ_HIP(hipGraphLaunch(c->gexec_long, c->stream));
  for (; k + kGraphSteps <= nsteps; k += kGraphSteps) LJ_HIP(hipGraphLaunch(c->gexec_run, c->stream));
  for (; k < nsteps; ++k) LJ_HIP(hipGraphLaunch(c->gexec, c->stream));
  return LJ_OK;
}

int compute_forces_internal(lj_ctx* c)
{
  hipStream_t st = c->stream;
  if (is_fp64(c)) {
    LJ_HIP(enqueue_force64(c, st));
    LJ_HIP(launch_reduce64(st, step64_args(c, 0.)));
  } else {
    LJ_TRY(enqueue_force(c, st, c->pos));
    LJ_HIP(launch_reduce_forces(st, fplan(c), c->ps, c->fpart, c->apart, c->force, c->blk));
  }
  LJ_HIP(launch_finalize_forces(st, c->blk, c->nblk, c->cfg.eps, c->ev, nullptr));
  // V and W into the observables block (CalculateForces sets V and P, MDSystem.cpp:257-309)
  LJ_HIP(launch_store_vw(st, c->ev, c->obs));
  return LJ_OK;
}

}  // namespace lj

extern "C" {

const char* lj_last_error(void) { return g_err.c_str(); }

int lj_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* lj_version(void) { return "ljhip 0.1 (gfx950)"; }

void lj_config_init(lj_config* cfg)
{
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->N = 128;  // MDSystem.h:32-39
  cfg->T0 = 1.5;
  cfg->rho = 0.2;
  cfg->canonical = LJ_EVN;
  cfg->boundary = LJ_BC_PERIODIC;
  cfg->eps = 1.0;
  cfg->sigma = 1.0;
  cfg->precision = LJ_PREC_FP32;
  cfg->jslices = 0;
  cfg->device = 0;
}

double lj_box_length(const lj_ctx* ctx) { return ctx ? ctx->L : 0.; }

float lj_rdf_dr2_default(int N) { return N > 0 ? rdf_dr2_rule(N) : 0.f; }

int lj_destroy(lj_ctx* c)
{
  if (!c) return LJ_OK;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  adj_free(c);
  dev_free(c->pos);
  dev_free(c->vel);
  dev_free(c->force);
  dev_free(c->tF);
  dev_free(c->q);
  dev_free(c->fpart);
  dev_free(c->apart);
  dev_free(c->blk);
  dev_free(c->obs);
  dev_free(c->hist);
  dev_free(c->hist_base);
  dev_free(c->ev);
  dev_free(c->rdf);
  dev_free(c->edges2);
  dev_free(c->n3items);
  dev_free(c->step_ew);
  dev_free(c->vh);
  dev_free(c->shear);
  dev_free(c->sh_fpart);
  dev_free(c->sh_apart);
  dev_free(c->sh_blk);
  dev_free(c->sh_ev);
  dev_free(c->xd);
  dev_free(c->vd);
  dev_free(c->fd);
  dev_free(c->tfd);
  dev_free(c->part64);
  dev_free(c->apart64);
  fl_free(c->fl);
  dev_free(c->rdf_step);
  dev_free(c->step_tab);
  rdf_cells_free(c->rdf_cells);
  if (c->e0) (void)hipEventDestroy(c->e0);
  if (c->e1) (void)hipEventDestroy(c->e1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LJ_OK;
}

int lj_create(const lj_config* cfg, lj_ctx** out)
{
  if (!cfg || !out) return fail(LJ_ERR_ARG, "lj_create: null argument");
  *out = nullptr;
  if (cfg->N <= 0) return fail(LJ_ERR_ARG, "lj_create: N must be positive (got %d)", cfg->N);
  if (cfg->N > (1 << 26)) return fail(LJ_ERR_ARG, "lj_create: N too large (%d)", cfg->N);
  if (!(cfg->rho > 0.)) return fail(LJ_ERR_ARG, "lj_create: rho must be positive");
  if (!(cfg->T0 >= 0.)) return fail(LJ_ERR_ARG, "lj_create: T0 must be non-negative");
  if (cfg->boundary < 0 || cfg->boundary > 2) return fail(LJ_ERR_ARG, "lj_create: boundary must be 0, 1 or 2");
  if (cfg->canonical != 0 && cfg->canonical != 1) return fail(LJ_ERR_ARG, "lj_create: canonical must be 0 or 1");
  if (!(cfg->eps > 0.) || !(cfg->sigma > 0.)) return fail(LJ_ERR_ARG, "lj_create: eps and sigma must be positive");
  if (cfg->precision != LJ_PREC_FP32 && cfg->precision != LJ_PREC_COMPENSATED &&
      cfg->precision != LJ_PREC_FP64)
    return fail(LJ_ERR_ARG, "lj_create: unknown precision mode %d", cfg->precision);
  LJ_TRY(check_device());
  int ndev = 0;
  LJ_HIP(hipGetDeviceCount(&ndev));
  if (cfg->device < 0 || cfg->device >= ndev) return fail(LJ_ERR_ARG, "lj_create: no device %d", cfg->device);
  LJ_HIP(hipSetDevice(cfg->device));

  lj_ctx* c = new lj_ctx();
  c->cfg = *cfg;
  c->L = box_length(cfg->N, cfg->rho);
  refresh_scalars(c);
  c->plan = make_force_plan(cfg->N, cfg->jslices, 1);
  if (const char* e = std::getenv("LJ_EMULATE_SHARDS")) c->emulate_shards = std::max(1, std::atoi(e));
  // the pair-once kernel serves the FP32 and compensated modes at sizes that
  // fill the GPU; an explicit j-slice count keeps the sliced kernel
  if (cfg->jslices == 0 && n3_eligible(cfg->N)) {
    if (cfg->precision == LJ_PREC_FP32) c->plan3 = make_n3_plan(cfg->N, false);
    else if (cfg->precision == LJ_PREC_COMPENSATED && n3_comp_enabled())
      c->plan3 = make_n3_plan(cfg->N, true);
  }
  const size_t npart = (size_t)std::max(c->plan.S, c->plan3.S) * c->plan.Npad;
  const int N = cfg->N;
  c->nblk = (N + kBlock - 1) / kBlock;
  int rc = LJ_OK;
  auto A = [&](int r) { if (rc == LJ_OK) rc = r; };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    c->stream = nullptr;
    A(fail(LJ_ERR_DEVICE, "hipStreamCreate failed"));
  }
  A(dev_alloc(&c->pos, N));
  A(dev_alloc(&c->vel, N));
  A(dev_alloc(&c->force, N));
  A(dev_alloc(&c->tF, N));
  A(dev_alloc(&c->q, N));
  A(dev_alloc(&c->fpart, npart));
  A(dev_alloc(&c->apart, npart));
  if (c->plan3.n3) {
    // the full variant's items, then the time step's
    std::vector<int2> it((size_t)c->plan3.nitems + c->plan3.nitems_step);
    n3_items(c->plan3.ng, it.data());
    n3_step_items(c->plan3, it.data() + c->plan3.nitems);
    A(dev_alloc(&c->n3items, it.size()));
    // copies on the context's stream: a null-stream copy can break another
    // thread's stream capture
    if (rc == LJ_OK &&
        (hipMemcpyAsync(c->n3items, it.data(), sizeof(int2) * it.size(), hipMemcpyHostToDevice,
                        c->stream) != hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess))
      A(fail(LJ_ERR_DEVICE, "item table upload failed"));
    c->plan3.items = c->n3items;
    c->plan3.items_step = c->n3items + c->plan3.nitems;
    A(dev_alloc(&c->step_ew, (size_t)4 * c->plan3.step_wgs()));  // two buffers
  }
  // rows: nblk for the per-block kernels, step_rows(N) >= nblk for the step
  A(dev_alloc(&c->blk, 2 * blk_rows(c) * kBlkFields));
  A(dev_alloc(&c->obs, 1));
  A(dev_alloc(&c->hist_base, 1));
  A(dev_alloc(&c->ev, 2));
  if (is_fp64(c)) {
    A(dev_alloc(&c->xd, N));
    A(dev_alloc(&c->vd, N));
    A(dev_alloc(&c->fd, N));
    A(dev_alloc(&c->tfd, N));
    A(dev_alloc(&c->part64, (size_t)c->plan.S * c->plan.Npad));
    A(dev_alloc(&c->apart64, (size_t)c->plan.S * c->plan.Npad));
  }
  if (rc == LJ_OK && hipEventCreate(&c->e0) != hipSuccess) A(fail(LJ_ERR_DEVICE, "hipEventCreate"));
  if (rc == LJ_OK && hipEventCreate(&c->e1) != hipSuccess) A(fail(LJ_ERR_DEVICE, "hipEventCreate"));
  if (rc == LJ_OK) A(ensure_hist(c, 1024));
  if (rc == LJ_OK) {
    // zero state: host buffers copied so that no memset node ever enters a
    // graph, on the context's stream (see the item table above)
    std::vector<float> z((size_t)4 * N, 0.f);
    lj_obs o{};
    const size_t bytes = sizeof(float) * 4 * (size_t)N;
    hipStream_t st = c->stream;
    const auto up = [st](void* d, const void* h, size_t n) {
      return hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    if (!up(c->pos, z.data(), bytes) || !up(c->vel, z.data(), bytes) ||
        !up(c->force, z.data(), bytes) || !up(c->tF, z.data(), bytes) ||
        !up(c->obs, &o, sizeof(o)) || hipStreamSynchronize(st) != hipSuccess)
      A(fail(LJ_ERR_DEVICE, "initial upload failed"));
    if (rc == LJ_OK && is_fp64(c)) {
      std::vector<double> zd((size_t)4 * N, 0.);
      const size_t db = sizeof(double) * 4 * (size_t)N;
      if (!up(c->xd, zd.data(), db) || !up(c->vd, zd.data(), db) || !up(c->fd, zd.data(), db) ||
          !up(c->tfd, zd.data(), db) || hipStreamSynchronize(st) != hipSuccess)
        A(fail(LJ_ERR_DEVICE, "initial upload failed"));
    }
  }
  if (rc != LJ_OK) {
    std::string msg = g_err;
    lj_destroy(c);
    g_err = msg;
    return rc;
  }
  *out = c;
  return LJ_OK;
}

int lj_set_array(lj_ctx* c, int type, const float* host)
{
  LJ_TRY(ctx_setdev(c));
  if (!host) return fail(LJ_ERR_ARG, "lj_set_array: null host pointer");
  float4* d = type == LJ_ARRAY_POS ? c->pos : type == LJ_ARRAY_VEL ? c->vel : type == LJ_ARRAY_FORCE ? c->force : nullptr;
  if (!d) return fail(LJ_ERR_ARG, "lj_set_array: unknown array type %d", type);
  LJ_HIP(hipMemcpyAsync(d, host, sizeof(float4) * c->cfg.N, hipMemcpyHostToDevice, c->stream));
  if (is_fp64(c)) LJ_HIP(launch_widen(c->stream, d, state64(c, type), c->cfg.N));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_set_array_f64(lj_ctx* c, int type, const double* host)
{
  LJ_TRY(ctx_setdev(c));
  if (!host) return fail(LJ_ERR_ARG, "lj_set_array_f64: null host pointer");
  float4* d = type == LJ_ARRAY_POS ? c->pos : type == LJ_ARRAY_VEL ? c->vel : type == LJ_ARRAY_FORCE ? c->force : nullptr;
  if (!d) return fail(LJ_ERR_ARG, "lj_set_array_f64: unknown array type %d", type);
  const int N = c->cfg.N;
  if (is_fp64(c)) {
    double4* dd = state64(c, type);
    LJ_HIP(hipMemcpyAsync(dd, host, sizeof(double4) * N, hipMemcpyHostToDevice, c->stream));
    LJ_HIP(launch_narrow(c->stream, dd, d, N));
    LJ_HIP(hipStreamSynchronize(c->stream));
    return LJ_OK;
  }
  std::vector<float> f((size_t)4 * N);
  for (size_t k = 0; k < f.size(); ++k) f[k] = (float)host[k];
  return lj_set_array(c, type, f.data());
}

int lj_get_array_f64(lj_ctx* c, int type, double* host)
{
  LJ_TRY(ctx_setdev(c));
  if (!host) return fail(LJ_ERR_ARG, "lj_get_array_f64: null host pointer");
  if (type != LJ_ARRAY_POS && type != LJ_ARRAY_VEL && type != LJ_ARRAY_FORCE)
    return fail(LJ_ERR_ARG, "lj_get_array_f64: unknown array type %d", type);
  const int N = c->cfg.N;
  if (is_fp64(c)) {
    LJ_HIP(hipMemcpyAsync(host, state64(c, type), sizeof(double4) * N, hipMemcpyDeviceToHost,
                          c->stream));
    LJ_HIP(hipStreamSynchronize(c->stream));
    return LJ_OK;
  }
  std::vector<float> f((size_t)4 * N);
  LJ_TRY(lj_get_array(c, type, f.data()));
  for (size_t k = 0; k < f.size(); ++k) host[k] = f[k];
  return LJ_OK;
}

int lj_get_array(lj_ctx* c, int type, float* host)
{
  LJ_TRY(ctx_setdev(c));
  if (!host) return fail(LJ_ERR_ARG, "lj_get_array: null host pointer");
  float4* d = type == LJ_ARRAY_POS ? c->pos : type == LJ_ARRAY_VEL ? c->vel : type == LJ_ARRAY_FORCE ? c->force : nullptr;
  if (!d) return fail(LJ_ERR_ARG, "lj_get_array: unknown array type %d", type);
  LJ_HIP(hipMemcpyAsync(host, d, sizeof(float4) * c->cfg.N, hipMemcpyDeviceToHost, c->stream));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_get_observables(lj_ctx* c, lj_obs* out)
{
  LJ_TRY(ctx_setdev(c));
  if (!out) return fail(LJ_ERR_ARG, "lj_get_observables: null output");
  LJ_HIP(hipMemcpyAsync(out, c->obs, sizeof(lj_obs), hipMemcpyDeviceToHost, c->stream));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_compute_forces(lj_ctx* c, lj_obs* out)
{
  LJ_TRY(ctx_setdev(c));
  LJ_TRY(compute_forces_internal(c));
  if (out) return lj_get_observables(c, out);
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_compute_parameters(lj_ctx* c, lj_obs* out)
{
  LJ_TRY(ctx_setdev(c));
  const StepArgs a = step_args(c, 0.);
  if (is_fp64(c))
    LJ_HIP(launch_kinetic64(c->stream, c->vd, c->cfg.N, c->blk));
  else
    LJ_HIP(launch_kinetic_partials(c->stream, c->vel, c->cfg.N, c->blk));
  LJ_HIP(launch_params(c->stream, a, 0));
  if (out) return lj_get_observables(c, out);
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_reset_averaging(lj_ctx* c)
{
  LJ_TRY(ctx_setdev(c));
  LJ_HIP(launch_reset_averaging(c->stream, c->obs));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_sample_initial_state(lj_ctx* c, uint64_t seed)
{
  return lj_sample_initial_state_ex(c, seed, LJ_INIT_GAUSSIAN);
}

int lj_sample_initial_state_ex(lj_ctx* c, uint64_t seed, int velocities)
{
  LJ_TRY(ctx_setdev(c));
  if (velocities != LJ_INIT_GAUSSIAN && velocities != LJ_INIT_MAXWELL)
    return fail(LJ_ERR_ARG, "lj_sample_initial_state_ex: velocities must be LJ_INIT_GAUSSIAN or LJ_INIT_MAXWELL");
  const int N = c->cfg.N;
  std::vector<float> p((size_t)4 * N), v((size_t)4 * N);
  lattice_positions(N, c->L, p.data());
  if (velocities == LJ_INIT_MAXWELL) maxwell_velocities(N, c->cfg.T0, seed, v.data());
  else gaussian_velocities(N, c->cfg.T0, seed, v.data());
  LJ_TRY(lj_set_array(c, LJ_ARRAY_POS, p.data()));
  LJ_TRY(lj_set_array(c, LJ_ARRAY_VEL, v.data()));
  LJ_HIP(launch_reset_clock(c->stream, c->obs));
  // MDSystem::Reinitialize: CalculateForces, CalculateParameters, reset averages (:101-113)
  LJ_TRY(lj_compute_forces(c, nullptr));
  LJ_TRY(lj_compute_parameters(c, nullptr));
  return lj_reset_averaging(c);
}

int lj_integrate(lj_ctx* c, double dt, int nsteps, double* hist)
{
  LJ_TRY(ctx_setdev(c));
  if (nsteps < 0) return fail(LJ_ERR_ARG, "lj_integrate: nsteps < 0");
  if (!std::isfinite(dt)) return fail(LJ_ERR_ARG, "lj_integrate: dt not finite");
  if (nsteps == 0) return LJ_OK;
  LJ_TRY(ensure_hist(c, nsteps));
  LJ_TRY(ensure_graph(c, dt));
  hipStream_t st = c->stream;
  LJ_HIP(launch_set_hist_base(st, c->obs, c->hist_base));
  LJ_TRY(launch_steps(c, nsteps));
  c->rdf_step_n = c->smp_rdf == LJ_SAMPLE_RDF_LAST ? 1 : c->rdf_step_n + nsteps;
  if (hist)
    LJ_HIP(hipMemcpyAsync(hist, c->hist, sizeof(double) * 5 * nsteps, hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

int lj_integrate_async(lj_ctx* c, double dt, int nsteps)
{
  LJ_TRY(ctx_setdev(c));
  if (nsteps < 0) return fail(LJ_ERR_ARG, "lj_integrate_async: nsteps < 0");
  if (!std::isfinite(dt)) return fail(LJ_ERR_ARG, "lj_integrate_async: dt not finite");
  if (nsteps == 0) return LJ_OK;
  LJ_TRY(ensure_graph(c, dt));
  LJ_HIP(launch_set_hist_base(c->stream, c->obs, c->hist_base));
  LJ_TRY(launch_steps(c, nsteps));
  c->rdf_step_n = c->smp_rdf == LJ_SAMPLE_RDF_LAST ? 1 : c->rdf_step_n + nsteps;
  return LJ_OK;
}

}  // extern "C"

namespace lj {
// validates an RDF request, sizes the context's histogram and edge table and
// fills the kernel arguments; the histogram is cleared unless accumulating
// into the same binning
int prepare_rdf(lj_ctx* c, int rule, int nbins, float dr2, double rmax, int accumulate,
                const char* who, RdfArgs* out)
{
  if (rule < LJ_RDF_REF_CPU || rule > LJ_RDF_RBINS) return fail(LJ_ERR_ARG, "%s: unknown rule %d", who, rule);
  if (nbins <= 0 || nbins > 8000) return fail(LJ_ERR_ARG, "%s: nbins must be in [1, 8000]", who);
  if (!(dr2 > 0.f)) dr2 = rdf_dr2_rule(c->cfg.N);
  if (!(rmax > 0.)) rmax = c->L / 2.;
  hipStream_t st = c->stream;
  if (nbins > c->rdf_cap) {
    LJ_HIP(hipStreamSynchronize(st));
    dev_free(c->rdf);
    c->rdf_cap = 0;
    LJ_TRY(dev_alloc(&c->rdf, nbins));
    c->rdf_cap = nbins;
    accumulate = 0;
  }
  const bool same = rule == c->rdf_rule && nbins == c->rdf_nbins && dr2 == c->rdf_dr2 && rmax == c->rdf_rmax;
  if (!accumulate || !same) LJ_HIP(launch_fill_u64(st, c->rdf, nbins));
  c->rdf_rule = rule;
  c->rdf_nbins = nbins;
  c->rdf_dr2 = dr2;
  c->rdf_rmax = rmax;
  RdfArgs a{};
  a.pos = c->pos;
  a.N = c->cfg.N;
  a.periodic = c->cfg.boundary == LJ_BC_PERIODIC;
  a.L = c->L;
  a.rule = rule;
  a.nbins = nbins;
  a.hist = c->rdf;
  if (rule == LJ_RDF_REF_GPU) gpu_minimage_thresholds(c->L, a.thr);
  else cpu_minimage_thresholds(c->L, a.thr);
  a.est_scale = rule == LJ_RDF_RBINS ? (float)(nbins / rmax) : (float)(1.0 / dr2);
  std::vector<float> e(nbins + 1);
  rdf_bin_thresholds(rule, nbins, dr2, rmax, e.data());
  a.end = rule == LJ_RDF_REF_GPU ? e[nbins - 1] : e[nbins];
  LJ_TRY(attach_rdf_cells(c, a));
  if (nbins + 1 > c->edges_cap) {
    LJ_HIP(hipStreamSynchronize(st));
    dev_free(c->edges2);
    c->edges_cap = 0;
    LJ_TRY(dev_alloc(&c->edges2, nbins + 1));
    c->edges_cap = nbins + 1;
  }
  // synchronous: the host table goes out of scope on return
  LJ_HIP(hipMemcpyAsync(c->edges2, e.data(), sizeof(float) * (nbins + 1), hipMemcpyHostToDevice, st));
  LJ_HIP(hipStreamSynchronize(st));
  a.tab = c->edges2;
  *out = a;
  return LJ_OK;
}
}  // namespace lj

extern "C" {

int lj_rdf(lj_ctx* c, int rule, int nbins, float dr2, double rmax, uint64_t* counts,
           int accumulate)
{
  LJ_TRY(ctx_setdev(c));
  if (!counts) return fail(LJ_ERR_ARG, "lj_rdf: null counts");
  RdfArgs a{};
  LJ_TRY(prepare_rdf(c, rule, nbins, dr2, rmax, accumulate, "lj_rdf", &a));
  hipStream_t st = c->stream;
  LJ_HIP(launch_rdf(st, a));
  if (rdf_uses_cells(a)) {
    // coordinates beyond [-L, 2 L): the cell-list pass counted nothing; the
    // all-pairs pass counts instead
    unsigned bad = 0u;
    LJ_HIP(hipMemcpyAsync(&bad, a.cells->result, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    LJ_HIP(hipStreamSynchronize(st));
    if (bad) {
      LJ_HIP(hipMemsetAsync(a.cells->result, 0, sizeof(unsigned), st));
      RdfArgs b = a;
      b.cells = nullptr;
      LJ_HIP(launch_rdf(st, b));
    }
  }
  LJ_HIP(hipMemcpyAsync(counts, c->rdf, sizeof(uint64_t) * a.nbins, hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

int lj_time_rdf(lj_ctx* c, int rule, int nbins, float dr2, double rmax, int reps, double* us)
{
  LJ_TRY(ctx_setdev(c));
  if (reps <= 0 || !us) return fail(LJ_ERR_ARG, "lj_time_rdf: bad arguments");
  RdfArgs a{};
  LJ_TRY(prepare_rdf(c, rule, nbins, dr2, rmax, 0, "lj_time_rdf", &a));
  hipStream_t st = c->stream;
  LJ_HIP(launch_rdf(st, a));  // warm
  LJ_HIP(hipEventRecord(c->e0, st));
  for (int k = 0; k < reps; ++k) LJ_HIP(launch_rdf(st, a));
  LJ_HIP(hipEventRecord(c->e1, st));
  LJ_HIP(hipEventSynchronize(c->e1));
  float ms = 0.f;
  LJ_HIP(hipEventElapsedTime(&ms, c->e0, c->e1));
  *us = 1000.0 * ms / reps;
  // the histogram now holds reps + 1 passes: the next lj_rdf starts afresh
  c->rdf_rule = -1;
  return LJ_OK;
}

int lj_velocity_histogram(lj_ctx* c, double step, int nbins, uint64_t* counts)
{
  LJ_TRY(ctx_setdev(c));
  if (!(step > 0.) || !std::isfinite(step)) return fail(LJ_ERR_ARG, "lj_velocity_histogram: step must be positive");
  if (nbins <= 0 || nbins > 16384) return fail(LJ_ERR_ARG, "lj_velocity_histogram: nbins must be in [1, 16384]");
  if (!counts) return fail(LJ_ERR_ARG, "lj_velocity_histogram: null counts");
  hipStream_t st = c->stream;
  if (nbins > c->vh_cap) {
    LJ_HIP(hipStreamSynchronize(st));
    dev_free(c->vh);
    c->vh_cap = 0;
    LJ_TRY(dev_alloc(&c->vh, nbins));
    c->vh_cap = nbins;
  }
  LJ_HIP(launch_fill_u64(st, c->vh, nbins));
  LJ_HIP(launch_velocity_hist(st, c->vel, c->cfg.N, step, nbins, c->vh));
  LJ_HIP(hipMemcpyAsync(counts, c->vh, sizeof(uint64_t) * nbins, hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

namespace {
// the plan and workspace of one shard's force pass
int shard_setup(lj_ctx* c, int shard, int nshards, int jslices, const char* fn, ForcePlan* out)
{
  LJ_TRY(ctx_setdev(c));
  const int N = c->cfg.N;
  if (nshards < 1 || nshards > (N + 63) / 64 || shard < 0 || shard >= nshards)
    return fail(LJ_ERR_ARG, "%s: need 0 <= shard < nshards <= ceil(N / 64) (got %d of %d)", fn,
                shard, nshards);
  if (jslices < 0 || jslices > N) return fail(LJ_ERR_ARG, "%s: jslices must be in [0, N]", fn);
  if (is_fp64(c)) return fail(LJ_ERR_STATE, "%s: FP32 and compensated modes only", fn);
  const ForcePlan pl = make_shard_plan(N, shard, nshards, jslices);
  const size_t np = (size_t)pl.S * pl.Npad;
  const int nb = (pl.iend() - pl.i0 + kBlock - 1) / kBlock;
  if (np > c->sh_cap || nb > c->sh_blk_cap || !c->sh_ev) {
    // grown to at least the old capacities (shares of other sizes come next)
    const size_t cap = std::max(np, c->sh_cap);
    const int bcap = std::max(nb, c->sh_blk_cap);
    LJ_HIP(hipStreamSynchronize(c->stream));
    dev_free(c->sh_fpart);
    dev_free(c->sh_apart);
    dev_free(c->sh_blk);
    dev_free(c->sh_ev);
    c->sh_cap = 0;
    c->sh_blk_cap = 0;
    LJ_TRY(dev_alloc(&c->sh_fpart, cap));
    LJ_TRY(dev_alloc(&c->sh_apart, cap));
    LJ_TRY(dev_alloc(&c->sh_blk, (size_t)bcap * kBlkFields));
    LJ_TRY(dev_alloc(&c->sh_ev, 2));
    c->sh_cap = cap;
    c->sh_blk_cap = bcap;
  }
  if (c->ps.periodic) LJ_HIP(launch_prep_q(c->stream, c->pos, c->q, N, c->ps.qscale));
  *out = pl;
  return LJ_OK;
}
}  // namespace

int lj_forces_shard(lj_ctx* c, int shard, int nshards, int jslices, float* d_force4, double* ev)
{
  ForcePlan pl;
  LJ_TRY(shard_setup(c, shard, nshards, jslices, "lj_forces_shard", &pl));
  hipStream_t st = c->stream;
  const bool comp = c->cfg.precision == LJ_PREC_COMPENSATED;
  LJ_HIP(launch_force(st, pl, c->ps, c->q, c->pos, c->sh_fpart, c->sh_apart, comp, true));
  float4* out = d_force4 ? reinterpret_cast<float4*>(d_force4) : c->force;
  LJ_HIP(launch_reduce_forces(st, pl, c->ps, c->sh_fpart, c->sh_apart, out, c->sh_blk));
  LJ_HIP(launch_finalize_forces(st, c->sh_blk, (pl.iend() - pl.i0 + kBlock - 1) / kBlock,
                                c->cfg.eps, c->sh_ev, nullptr));
  if (ev) LJ_HIP(hipMemcpyAsync(ev, c->sh_ev, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

int lj_time_forces_shard(lj_ctx* c, int shard, int nshards, int jslices, int reps, double* us)
{
  if (reps <= 0 || !us) return fail(LJ_ERR_ARG, "lj_time_forces_shard: bad arguments");
  ForcePlan pl;
  LJ_TRY(shard_setup(c, shard, nshards, jslices, "lj_time_forces_shard", &pl));
  hipStream_t st = c->stream;
  const bool comp = c->cfg.precision == LJ_PREC_COMPENSATED;
  LJ_HIP(launch_force(st, pl, c->ps, c->q, c->pos, c->sh_fpart, c->sh_apart, comp, false));  // warm
  LJ_HIP(hipEventRecord(c->e0, st));
  for (int k = 0; k < reps; ++k)
    LJ_HIP(launch_force(st, pl, c->ps, c->q, c->pos, c->sh_fpart, c->sh_apart, comp, false));
  LJ_HIP(hipEventRecord(c->e1, st));
  LJ_HIP(hipEventSynchronize(c->e1));
  float ms = 0.f;
  LJ_HIP(hipEventElapsedTime(&ms, c->e0, c->e1));
  *us = 1000.0 * ms / reps;
  return LJ_OK;
}

int lj_time_force_items_share(lj_ctx* c, int shard, int nshards, int reps, double* us)
{
  LJ_TRY(ctx_setdev(c));
  const ForcePlan& pl = c->plan3;
  if (reps <= 0 || !us) return fail(LJ_ERR_ARG, "lj_time_force_items_share: bad arguments");
  if (pl.n3 != 2 || pl.split != 1 || pl.comp)
    return fail(LJ_ERR_STATE, "lj_time_force_items_share: needs the FP32 pair-once step pass "
                              "with paired diagonal items");
  const int M = pl.nitems_step;
  if (nshards < 1 || nshards > M || shard < 0 || shard >= nshards)
    return fail(LJ_ERR_ARG, "lj_time_force_items_share: need 0 <= shard < nshards <= %d", M);
  const int first = (int)((long)shard * M / nshards);
  const int count = (int)((long)(shard + 1) * M / nshards) - first;
  hipStream_t st = c->stream;
  if (c->ps.periodic) LJ_HIP(launch_prep_q(st, c->pos, c->q, c->cfg.N, c->ps.qscale));
  LJ_HIP(launch_force_n3_step_items(st, pl, c->ps, c->q, c->pos, c->fpart, c->step_ew, first,
                                    count));  // warm
  LJ_HIP(hipEventRecord(c->e0, st));
  for (int k = 0; k < reps; ++k)
    LJ_HIP(launch_force_n3_step_items(st, pl, c->ps, c->q, c->pos, c->fpart, c->step_ew, first,
                                      count));
  LJ_HIP(hipEventRecord(c->e1, st));
  LJ_HIP(hipEventSynchronize(c->e1));
  float ms = 0.f;
  LJ_HIP(hipEventElapsedTime(&ms, c->e0, c->e1));
  *us = 1000.0 * ms / reps;
  return LJ_OK;
}

int lj_shear_pressure(lj_ctx* c, double* pshear)
{
  LJ_TRY(ctx_setdev(c));
  if (!pshear) return fail(LJ_ERR_ARG, "lj_shear_pressure: null output");
  hipStream_t st = c->stream;
  const ForcePlan& pl = c->plan;  // the sliced grid, whatever kernel the steps run
  const int n = pl.gx * pl.S + 1;
  if (n > c->shear_cap) {
    LJ_HIP(hipStreamSynchronize(st));
    dev_free(c->shear);
    c->shear_cap = 0;
    LJ_TRY(dev_alloc(&c->shear, n));
    c->shear_cap = n;
  }
  // FP64 contexts: the float mirrors of the state
  if (c->ps.periodic) LJ_HIP(launch_prep_q(st, c->pos, c->q, c->cfg.N, c->ps.qscale));
  LJ_HIP(launch_shear(st, pl, c->ps, c->q, c->pos, c->vel, c->cfg.N / c->cfg.rho, c->shear,
                      c->shear + n - 1));
  LJ_HIP(hipMemcpyAsync(pshear, c->shear + n - 1, sizeof(double), hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

int lj_renormalize_velocities(lj_ctx* c, int recompute_tkin)
{
  LJ_TRY(ctx_setdev(c));
  double factor = 0.;
  if (!recompute_tkin) {
    // Tkin = T of the last CalculateParameters (MDSystem.cpp:377)
    lj_obs o;
    LJ_TRY(lj_get_observables(c, &o));
    if (!(o.T > 0.)) return fail(LJ_ERR_STATE, "lj_renormalize_velocities: T is not positive");
    factor = std::sqrt(c->cfg.T0 / o.T);
  } else if (is_fp64(c)) {
    LJ_HIP(launch_kinetic64(c->stream, c->vd, c->cfg.N, c->blk));
  } else {
    LJ_HIP(launch_kinetic_partials(c->stream, c->vel, c->cfg.N, c->blk));
  }
  if (is_fp64(c))
    LJ_HIP(launch_scale64(c->stream, c->vd, c->vel, c->cfg.N, c->blk, c->nblk, c->cfg.T0, factor,
                          c->obs));
  else
    LJ_HIP(launch_scale_velocities(c->stream, c->vel, c->cfg.N, c->blk, c->nblk, c->cfg.T0,
                                   factor, c->obs));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_renormalize_to_energy(lj_ctx* c, double ust)
{
  LJ_TRY(ctx_setdev(c));
  lj_obs o;
  LJ_TRY(lj_get_observables(c, &o));
  const double Kdes = ust * c->cfg.N - o.V;  // MDSystem.cpp:393-395
  if (!(Kdes > 0.) || !(o.K > 0.))
    return fail(LJ_ERR_ARG, "lj_renormalize_to_energy: target kinetic energy %g not reachable", Kdes);
  const double factor = std::sqrt(Kdes / o.K);
  if (is_fp64(c))
    LJ_HIP(launch_scale64(c->stream, c->vd, c->vel, c->cfg.N, c->blk, c->nblk, c->cfg.T0, factor,
                          nullptr));
  else
    LJ_HIP(launch_scale_velocities(c->stream, c->vel, c->cfg.N, c->blk, c->nblk, c->cfg.T0,
                                   factor, nullptr));
  o.K = Kdes;
  o.U = o.K + o.V;
  LJ_HIP(hipMemcpyAsync(c->obs, &o, sizeof(o), hipMemcpyHostToDevice, c->stream));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_apply_boundary_conditions(lj_ctx* c)
{
  LJ_TRY(ctx_setdev(c));
  if (is_fp64(c))
    LJ_HIP(launch_bc64(c->stream, c->xd, c->vd, c->pos, c->vel, c->cfg.N, c->cfg.boundary, c->L));
  else
    LJ_HIP(launch_apply_bc(c->stream, c->pos, c->vel, c->cfg.N, c->cfg.boundary, c->L));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_set_mode(lj_ctx* c, int canonical, int boundary)
{
  LJ_TRY(ctx_setdev(c));
  if (canonical != 0 && canonical != 1) return fail(LJ_ERR_ARG, "lj_set_mode: canonical must be 0 or 1");
  if (boundary < 0 || boundary > 2) return fail(LJ_ERR_ARG, "lj_set_mode: boundary must be 0, 1 or 2");
  // replays enqueued by lj_integrate_async may still run the graphs dropped here
  LJ_HIP(hipStreamSynchronize(c->stream));
  c->cfg.canonical = canonical;
  c->cfg.boundary = boundary;
  refresh_scalars(c);
  drop_graph(c);
  return LJ_OK;
}

int lj_synchronize(lj_ctx* c)
{
  LJ_TRY(ctx_setdev(c));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_device_arrays(lj_ctx* c, float** pos, float** vel, float** force)
{
  if (!c) return fail(LJ_ERR_ARG, "null context");
  if (pos) *pos = reinterpret_cast<float*>(c->pos);
  if (vel) *vel = reinterpret_cast<float*>(c->vel);
  if (force) *force = reinterpret_cast<float*>(c->force);
  return LJ_OK;
}

int lj_time_force_kernel_stamped(lj_ctx* c, int reps, double* us_kernel, double* us_span)
{
  LJ_TRY(ctx_setdev(c));
  if (reps <= 0 || reps > 4096 || !us_kernel || !us_span)
    return fail(LJ_ERR_ARG, "lj_time_force_kernel_stamped: bad arguments");
  if (is_fp64(c)) return fail(LJ_ERR_STATE, "lj_time_force_kernel_stamped: FP32 mode only");
  std::vector<hipEvent_t> ev((size_t)2 * reps, nullptr);
  struct Cleanup {  // the stamp is cleared and the events destroyed on every path
    std::vector<hipEvent_t>& ev;
    ~Cleanup()
    {
      lj::tl_force_stamp = lj::LaunchStamp{};
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    }
  } cleanup{ev};
  for (hipEvent_t& e : ev) LJ_HIP(hipEventCreate(&e));
  hipStream_t st = c->stream;
  if (c->ps.periodic) LJ_HIP(launch_prep_q(st, c->pos, c->q, c->cfg.N, c->ps.qscale));
  const bool comp = c->cfg.precision == LJ_PREC_COMPENSATED;
  LJ_HIP(launch_force(st, fplan(c), c->ps, c->q, c->pos, c->fpart, c->apart, comp, false,
                      c->step_ew));  // warm
  for (int k = 0; k < reps; ++k) {
    lj::tl_force_stamp = lj::LaunchStamp{ev[2 * k], ev[2 * k + 1]};
    LJ_HIP(launch_force(st, fplan(c), c->ps, c->q, c->pos, c->fpart, c->apart, comp, false,
                        c->step_ew));
  }
  lj::tl_force_stamp = lj::LaunchStamp{};
  LJ_HIP(hipEventSynchronize(ev.back()));
  double sum = 0.0;
  for (int k = 0; k < reps; ++k) {
    float ms = 0.f;
    LJ_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
    sum += ms;
  }
  float span = 0.f;
  LJ_HIP(hipEventElapsedTime(&span, ev[0], ev.back()));
  *us_kernel = 1000.0 * sum / reps;
  *us_span = 1000.0 * span / reps;
  return LJ_OK;
}

int lj_time_force_kernel(lj_ctx* c, int reps, double* us)
{
  LJ_TRY(ctx_setdev(c));
  if (reps <= 0 || !us) return fail(LJ_ERR_ARG, "lj_time_force_kernel: bad arguments");
  hipStream_t st = c->stream;
  if (is_fp64(c)) {
    LJ_HIP(enqueue_force64(c, st));  // warm
    LJ_HIP(hipEventRecord(c->e0, st));
    for (int k = 0; k < reps; ++k) LJ_HIP(enqueue_force64(c, st));
    LJ_HIP(hipEventRecord(c->e1, st));
    LJ_HIP(hipEventSynchronize(c->e1));
    float ms = 0.f;
    LJ_HIP(hipEventElapsedTime(&ms, c->e0, c->e1));
    *us = 1000.0 * ms / reps;
    return LJ_OK;
  }
  if (c->ps.periodic) LJ_HIP(launch_prep_q(st, c->pos, c->q, c->cfg.N, c->ps.qscale));
  const bool comp = c->cfg.precision == LJ_PREC_COMPENSATED;
  // the force pass of the time step (force-only partials with the pair-once kernel)
  LJ_HIP(launch_force(st, fplan(c), c->ps, c->q, c->pos, c->fpart, c->apart, comp, false,
                      c->step_ew));  // warm
  LJ_HIP(hipEventRecord(c->e0, st));
  for (int k = 0; k < reps; ++k)
    LJ_HIP(launch_force(st, fplan(c), c->ps, c->q, c->pos, c->fpart, c->apart, comp, false,
                        c->step_ew));
  LJ_HIP(hipEventRecord(c->e1, st));
  LJ_HIP(hipEventSynchronize(c->e1));
  float ms = 0.f;
  LJ_HIP(hipEventElapsedTime(&ms, c->e0, c->e1));
  *us = 1000.0 * ms / reps;
  return LJ_OK;
}

int lj_force_plan_info(lj_ctx* c, int* pair_once, int* nslots)
{
  if (!c) return fail(LJ_ERR_ARG, "null context");
  const ForcePlan& pl = is_fp64(c) ? c->plan : fplan(c);
  if (pair_once) *pair_once = pl.n3;
  if (nslots) *nslots = pl.S;
  return LJ_OK;
}

int lj_forces_device(lj_ctx* c, const float* d_pos4, float* d_force4, double* d_ev, void* stream)
{
  LJ_TRY(ctx_setdev(c));
  if (!d_pos4 || !d_force4) return fail(LJ_ERR_ARG, "lj_forces_device: null device pointer");
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  const float4* p = reinterpret_cast<const float4*>(d_pos4);
  LJ_TRY(enqueue_force(c, st, p));
  LJ_HIP(launch_reduce_forces(st, fplan(c), c->ps, c->fpart, c->apart,
                              reinterpret_cast<float4*>(d_force4), c->blk));
  LJ_HIP(launch_finalize_forces(st, c->blk, c->nblk, c->cfg.eps, d_ev ? d_ev : c->ev, nullptr));
  return LJ_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// compatibility set (MDSystem.cu:155-299)

namespace {

struct CompatScratch {
  int device = -1;
  hipStream_t st = nullptr;
  int capN = 0;
  size_t capPart = 0;
  uint4* q = nullptr;
  float4* fpart = nullptr;
  float* apart = nullptr;
  double* blk = nullptr;
  unsigned long long* rdf = nullptr;
  float* pressure = nullptr;
  float* rdf_tab = nullptr;  // bin thresholds of the GPU rule for tab_dr2
  float tab_dr2 = 0.f, tab_last = 0.f;
};

std::mutex g_compat_mu;
std::vector<CompatScratch> g_compat;

void compat_report(int rc)
{
  if (rc == LJ_OK) return;
  std::fprintf(stderr, "ljhip: %s\n", g_err.c_str());
  const char* ab = std::getenv("LJ_COMPAT_ABORT");
  if (ab && ab[0] == '1') std::abort();
}

void compat_release(CompatScratch& s)
{
  if (s.device >= 0) (void)hipSetDevice(s.device);
  if (s.st) (void)hipStreamSynchronize(s.st);
  dev_free(s.q);
  dev_free(s.fpart);
  dev_free(s.apart);
  dev_free(s.blk);
  dev_free(s.rdf);
  dev_free(s.pressure);
  dev_free(s.rdf_tab);
  if (s.st) (void)hipStreamDestroy(s.st);
  s = CompatScratch();
}

int compat_scratch(int N, const ForcePlan& pl, CompatScratch** out)
{
  int dev = 0;
  LJ_HIP(hipGetDevice(&dev));
  if ((int)g_compat.size() <= dev) g_compat.resize(dev + 1);
  CompatScratch& s = g_compat[dev];
  s.device = dev;
  if (!s.st) LJ_HIP(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
  const size_t part = (size_t)pl.S * pl.Npad;
  if (N > s.capN || part > s.capPart) {
    LJ_HIP(hipStreamSynchronize(s.st));
    dev_free(s.q);
    dev_free(s.fpart);
    dev_free(s.apart);
    dev_free(s.blk);
    s.capN = 0;
    LJ_TRY(dev_alloc(&s.q, N));
    LJ_TRY(dev_alloc(&s.fpart, (size_t)pl.S * pl.Npad));
    LJ_TRY(dev_alloc(&s.apart, (size_t)pl.S * pl.Npad));
    LJ_TRY(dev_alloc(&s.blk, (size_t)((N + kBlock - 1) / kBlock) * kBlkFields));
    s.capN = N;
    s.capPart = part;
  }
  if (!s.rdf) LJ_TRY(dev_alloc(&s.rdf, 256));
  if (!s.rdf_tab) LJ_TRY(dev_alloc(&s.rdf_tab, 257));
  if (!s.pressure) LJ_TRY(dev_alloc(&s.pressure, 1));
  *out = &s;
  return LJ_OK;
}

int compat_forces(float* Pos, float* Force, float* host_pressure, int N, float host_L,
                  int Lperiodic, int* host_RDF, float host_dr2)
{
  if (!Pos || !Force) return fail(LJ_ERR_ARG, "calculateNForces: null device array");
  if (N <= 0) return fail(LJ_ERR_ARG, "calculateNForces: numBodies must be positive (got %d)", N);
  if (Lperiodic && !(host_L > 0.f)) return fail(LJ_ERR_ARG, "calculateNForces: box side must be positive");
  if (host_RDF && !(host_dr2 > 0.f)) return fail(LJ_ERR_ARG, "calculateNForces: RDF bin width must be positive");
  LJ_TRY(check_device());
  const ForcePlan pl = make_force_plan(N, 0, 1);
  const PairScalars ps = make_pair_scalars((double)host_L, Lperiodic != 0, 1., 1., true);
  CompatScratch* s = nullptr;
  LJ_TRY(compat_scratch(N, pl, &s));
  hipStream_t st = s->st;
  const float4* p = reinterpret_cast<const float4*>(Pos);
  if (ps.periodic) LJ_HIP(launch_prep_q(st, p, s->q, N, ps.qscale));
  LJ_HIP(launch_force(st, pl, ps, s->q, p, s->fpart, s->apart, false, true));
  LJ_HIP(launch_reduce_forces(st, pl, ps, s->fpart, s->apart, reinterpret_cast<float4*>(Force), s->blk));
  LJ_HIP(launch_finalize_forces(st, s->blk, (N + kBlock - 1) / kBlock, 1., nullptr, s->pressure));
  if (host_RDF) {
    LJ_HIP(launch_fill_u64(st, s->rdf, 256));
    if (s->tab_dr2 != host_dr2) {
      float tab[257];
      rdf_bin_thresholds(LJ_RDF_REF_GPU, 256, host_dr2, 0., tab);
      LJ_HIP(hipMemcpyAsync(s->rdf_tab, tab, sizeof(tab), hipMemcpyHostToDevice, st));
      LJ_HIP(hipStreamSynchronize(st));  // the host table goes out of scope
      s->tab_dr2 = host_dr2;
      s->tab_last = tab[255];
    }
    RdfArgs a{};
    a.pos = p;
    a.N = N;
    a.periodic = Lperiodic != 0;
    a.L = (double)host_L;
    a.rule = LJ_RDF_REF_GPU;
    a.nbins = 256;
    a.tab = s->rdf_tab;
    a.est_scale = 1.f / host_dr2;
    gpu_minimage_thresholds((double)host_L, a.thr);
    a.end = s->tab_last;
    a.hist = s->rdf;
    LJ_HIP(launch_rdf(st, a));
  }
  float pr = 0.f;
  unsigned long long h[256];
  LJ_HIP(hipMemcpyAsync(&pr, s->pressure, sizeof(float), hipMemcpyDeviceToHost, st));
  if (host_RDF) LJ_HIP(hipMemcpyAsync(h, s->rdf, sizeof(h), hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  if (host_pressure) *host_pressure = pr;
  if (host_RDF)
    for (int b = 0; b < 256; ++b) host_RDF[b] = (int)h[b];
  return LJ_OK;
}

}  // namespace

extern "C" {

void allocateNBodyArrays(float* vel[2], int numBodies)
{
  if (!vel) {
    compat_report(fail(LJ_ERR_ARG, "allocateNBodyArrays: null argument"));
    return;
  }
  allocateArray(&vel[0], numBodies);
  allocateArray(&vel[1], numBodies);
}

void allocateArray(float** dest, int number)
{
  if (!dest || number < 0) {
    compat_report(fail(LJ_ERR_ARG, "allocateArray: bad argument"));
    return;
  }
  float4* p = nullptr;
  int rc = check_device();
  if (rc == LJ_OK) rc = dev_alloc(&p, (size_t)number);
  *dest = reinterpret_cast<float*>(p);
  compat_report(rc);
}

void deleteNBodyArrays(float* vel[2])
{
  if (!vel) return;
  deleteArray(vel[0]);
  deleteArray(vel[1]);
}

void deleteArray(float* arr)
{
  if (!arr) return;
  hipError_t e = hipFree(arr);
  if (e != hipSuccess) compat_report(fail(LJ_ERR_DEVICE, "deleteArray: %s", hipGetErrorString(e)));
}

void allocateIntArray(int** dest, int number)
{
  if (!dest || number < 0) {
    compat_report(fail(LJ_ERR_ARG, "allocateIntArray: bad argument"));
    return;
  }
  int* p = nullptr;
  int rc = check_device();
  if (rc == LJ_OK) rc = dev_alloc(&p, (size_t)number);
  *dest = p;
  compat_report(rc);
}

void deleteIntArray(int* arr)
{
  if (!arr) return;
  hipError_t e = hipFree(arr);
  if (e != hipSuccess) compat_report(fail(LJ_ERR_DEVICE, "deleteIntArray: %s", hipGetErrorString(e)));
}

void copyArrayFromDevice(float* host, const float* device, unsigned int /*pbo*/, int numBodies)
{
  if (!host || !device || numBodies < 0) {
    compat_report(fail(LJ_ERR_ARG, "copyArrayFromDevice: bad argument"));
    return;
  }
  hipError_t e = hipMemcpy(host, device, sizeof(float) * 4 * (size_t)numBodies, hipMemcpyDeviceToHost);
  if (e != hipSuccess) compat_report(fail(LJ_ERR_DEVICE, "copyArrayFromDevice: %s", hipGetErrorString(e)));
}

void copyArrayToDevice(float* device, const float* host, int numBodies)
{
  if (!host || !device || numBodies < 0) {
    compat_report(fail(LJ_ERR_ARG, "copyArrayToDevice: bad argument"));
    return;
  }
  hipError_t e = hipMemcpy(device, host, sizeof(float) * 4 * (size_t)numBodies, hipMemcpyHostToDevice);
  if (e != hipSuccess) compat_report(fail(LJ_ERR_DEVICE, "copyArrayToDevice: %s", hipGetErrorString(e)));
}

void registerGLBufferObject(unsigned int) {}
void unregisterGLBufferObject(unsigned int) {}

void threadSync(void)
{
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) compat_report(fail(LJ_ERR_DEVICE, "threadSync: %s", hipGetErrorString(e)));
}

void calculateNForces(float* Pos, float* Force, float* host_pressure, int numBodies, float host_L,
                      int Lperiodic, int* host_RDF, float host_dr2, int /*blockSize*/, int /*q*/)
{
  std::lock_guard<std::mutex> lk(g_compat_mu);
  compat_report(compat_forces(Pos, Force, host_pressure, numBodies, host_L, Lperiodic, host_RDF, host_dr2));
}

void threadExit(void)
{
  std::lock_guard<std::mutex> lk(g_compat_mu);
  for (auto& s : g_compat) compat_release(s);
  g_compat.clear();
}

}  // extern "C"
