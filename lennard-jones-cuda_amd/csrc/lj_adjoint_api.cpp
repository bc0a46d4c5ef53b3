// lj_adjoint_api.cpp -- differentiable trajectories: loss of M velocity-Verlet
// (EVN) steps and its gradient with respect to eps, sigma and the initial
// positions, by a reverse sweep over stored states with checkpoints every k
// steps and recomputation inside each segment.
//
// The differentiable trajectory keeps positions, velocities and forces in
// FP64 (pair arithmetic stays the FP32 force kernel on fixed-point
// coordinates): FP32 positions of the reference layout round each step's
// displacement (|x| ~ 13, ulp ~ 1e-6, dt v ~ 5e-3), and in a dense liquid the
// chaotic dynamics amplify that rounding far beyond the gradient tolerance.
//
// Loss over the states after steps m = 1..M (observables of
// MDSystem::CalculateParameters, MDSystem.cpp:325-359):
//   L = (1/M) sum_m E_pot(x_m) + w_P ((1/M) sum_m P_m - P_target)^2,
//   P_m = (W(x_m)/3 + sum_i |v_i|^2/3) / vol.
// Adjoint of one step x' = x + dt v + dt^2/2 F(x), v' = v + dt/2 (F(x) + F(x')):
//   u_m   = lx_{m+1} + dt/2 H(x_{m+1}) lv_{m+1}
//   w_m   = dt^2/2 u_m + dt/2 lv_{m+1}
//   lx_m  = u_m + H(x_m) w_m + g_x(m),   lv_m = dt u_m + lv_{m+1} + g_v(m)
// so that u_{m-1} = u_m + g_x(m) + H(x_m) z_m with z_m = w_m + dt/2 lv_m: one
// Hessian-vector product per stored state, and
//   dL/dtheta = sum_m (dF/dtheta(x_m))^T z_m + direct terms.
#include <cmath>
#include <cstring>
#include <vector>

#include "lj_adjoint.h"
#include "lj_ctx.h"
#include "lj_fp64.h"

using namespace lj;

struct AdjState {
  int N = 0, nblk = 0;
  int seg_cap = 0, ck_cap = 0;
  ForcePlan plan;  // pair pass geometry (adj_force_plan)
  float4* part = nullptr;
  float4* z = nullptr;
  double4 *hz = nullptr, *f = nullptr, *gw = nullptr, *u = nullptr, *lv = nullptr, *dx0 = nullptr;
  double* blk = nullptr;
  double* rowacc = nullptr;  // block partials added over the reverse sweep's states
  double* scal = nullptr;
  double* theta = nullptr;
  double* acc = nullptr;                                // {sum E, sum P, E, P}
  double4 *xd = nullptr, *vd = nullptr, *fd = nullptr;  // FP64 working state
  float4* xf = nullptr;                                 // float coordinates (open box)
  double4* ck = nullptr;   // checkpoints [ncheck][3][N] (x, v, f)
  double4* seg = nullptr;  // segment states [k+1][2][N] (x, v)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double g_dt = 0., g_eps = 0., g_sigma = 0.;
  int g_bc = -1;  // boundary the captured step was built for (periodic flag, image scale)
  // runs of g_k steps captured together (same key as the one-step graph):
  // the forward's, and the recompute's that also stores each state in its
  // segment slot 1..g_k
  hipGraph_t graph_k = nullptr, graph_seg = nullptr;
  hipGraphExec_t gexec_k = nullptr, gexec_seg = nullptr;
  int g_k = 0;
};

namespace {

void adj_drop_kgraphs(AdjState* a)
{
  if (a->gexec_k) (void)hipGraphExecDestroy(a->gexec_k);
  if (a->graph_k) (void)hipGraphDestroy(a->graph_k);
  if (a->gexec_seg) (void)hipGraphExecDestroy(a->gexec_seg);
  if (a->graph_seg) (void)hipGraphDestroy(a->graph_seg);
  a->gexec_k = a->gexec_seg = nullptr;
  a->graph_k = a->graph_seg = nullptr;
  a->g_k = 0;
}

void adj_drop_graph(AdjState* a)
{
  if (a->gexec) (void)hipGraphExecDestroy(a->gexec);
  if (a->graph) (void)hipGraphDestroy(a->graph);
  a->gexec = nullptr;
  a->graph = nullptr;
  adj_drop_kgraphs(a);
}

int adj_ensure(lj_ctx* c, int ncheck, int seg)
{
  AdjState*& a = c->adj;
  if (!a) {
    a = new AdjState();
    a->N = c->cfg.N;
    a->nblk = adj_rows(c->cfg.N);  // rows of k_adj_reduce / k_dkick
    const int N = a->N;
    a->plan = adj_force_plan(N);
    LJ_TRY(dev_alloc(&a->part, (size_t)3 * a->plan.S * a->plan.Npad));
    LJ_TRY(dev_alloc(&a->z, N));
    LJ_TRY(dev_alloc(&a->hz, N));
    LJ_TRY(dev_alloc(&a->f, N));
    LJ_TRY(dev_alloc(&a->gw, N));
    LJ_TRY(dev_alloc(&a->u, N));
    LJ_TRY(dev_alloc(&a->lv, N));
    LJ_TRY(dev_alloc(&a->dx0, N));
    LJ_TRY(dev_alloc(&a->blk, (size_t)a->nblk * 8));
    LJ_TRY(dev_alloc(&a->rowacc, (size_t)a->nblk * 8));
    LJ_TRY(dev_alloc(&a->scal, kAdjScalars));
    LJ_TRY(dev_alloc(&a->theta, 2));
    LJ_TRY(dev_alloc(&a->acc, 4));
    LJ_TRY(dev_alloc(&a->xd, N));
    LJ_TRY(dev_alloc(&a->vd, N));
    LJ_TRY(dev_alloc(&a->fd, N));
    LJ_TRY(dev_alloc(&a->xf, N));
  }
  if (ncheck > a->ck_cap) {
    LJ_HIP(hipStreamSynchronize(c->stream));
    dev_free(a->ck);
    a->ck_cap = 0;
    LJ_TRY(dev_alloc(&a->ck, (size_t)3 * ncheck * a->N));
    a->ck_cap = ncheck;
  }
  if (seg > a->seg_cap) {
    LJ_HIP(hipStreamSynchronize(c->stream));
    adj_drop_kgraphs(a);  // the recompute graph holds the old slots
    dev_free(a->seg);
    a->seg_cap = 0;
    LJ_TRY(dev_alloc(&a->seg, (size_t)2 * seg * a->N));
    a->seg_cap = seg;
  }
  return LJ_OK;
}

int dcopy(lj_ctx* c, double4* dst, const double4* src)
{
  LJ_HIP(hipMemcpyAsync(dst, src, sizeof(double4) * c->cfg.N, hipMemcpyDeviceToDevice, c->stream));
  return LJ_OK;
}

double qscale_of(lj_ctx* c) { return c->ps.periodic ? c->ps.qscale : 0.; }

// forces of the FP64 state into fd (and the block partials); ties resolved
int dforces(lj_ctx* c, hipStream_t st, const double4* x)
{
  AdjState* a = c->adj;
  LJ_HIP(launch_dprep(st, x, c->cfg.N, qscale_of(c), c->q, a->xf));
  LJ_HIP(launch_force(st, fplan(c), c->ps, c->q, a->xf, c->fpart, c->apart, false, true));
  LJ_HIP(launch_dkick(st, fplan(c), c->ps.FS, 0., c->fpart, c->apart, a->fd, a->vd, a->blk));
  return LJ_OK;
}

// one FP64 step: drift, forces, kick, observables accumulated into acc.
// slot (the recompute, may be null): the state goes to slot (x) and slot + N
// (v) from the kick kernel, and the observables, which only the forward
// sweep's loss reads, are left out
int enqueue_dstep(lj_ctx* c, hipStream_t st, double dt, double4* slot = nullptr)
{
  AdjState* a = c->adj;
  const double vol = std::pow(c->L, 3.);
  const int N = c->cfg.N;
  LJ_HIP(launch_ddrift(st, a->xd, a->vd, a->fd, N, dt, qscale_of(c), c->q, a->xf));
  LJ_HIP(launch_force(st, fplan(c), c->ps, c->q, a->xf, c->fpart, c->apart, false, false));
  LJ_HIP(launch_dkick(st, fplan(c), c->ps.FS, dt, c->fpart, c->apart, a->fd, a->vd, a->blk, a->xd,
                      slot, slot ? slot + N : nullptr));
  if (!slot) LJ_HIP(launch_dfinal(st, a->blk, a->nblk, c->cfg.eps, vol, 1, a->acc));
  return LJ_OK;
}

int ensure_dgraph(lj_ctx* c, double dt)
{
  AdjState* a = c->adj;
  if (a->gexec && a->g_dt == dt && a->g_eps == c->cfg.eps && a->g_sigma == c->cfg.sigma &&
      a->g_bc == c->cfg.boundary)
    return LJ_OK;
  adj_drop_graph(a);
  hipStream_t st = c->stream;
  LJ_HIP(hipStreamSynchronize(st));
  LJ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_dstep(c, st, dt);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(st, &g);
  if (rc != LJ_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  LJ_HIP(ee);
  a->graph = g;
  LJ_HIP(hipGraphInstantiate(&a->gexec, a->graph, nullptr, nullptr, 0));
  a->g_dt = dt;
  a->g_eps = c->cfg.eps;
  a->g_sigma = c->cfg.sigma;
  a->g_bc = c->cfg.boundary;
  return LJ_OK;
}

// captures `steps` FP64 steps back to back; with_slots: step s also writes
// its state into segment slot s (1..steps), without the loss observables
int capture_dsteps(lj_ctx* c, double dt, int steps, bool with_slots, hipGraph_t* graph,
                   hipGraphExec_t* exec)
{
  AdjState* a = c->adj;
  const int N = c->cfg.N;
  hipStream_t st = c->stream;
  LJ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = LJ_OK;
  for (int s = 1; s <= steps && rc == LJ_OK; ++s)
    rc = enqueue_dstep(c, st, dt, with_slots ? a->seg + (size_t)2 * s * N : nullptr);
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(st, &g);
  if (rc != LJ_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  LJ_HIP(ee);
  *graph = g;
  LJ_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return LJ_OK;
}

// the k-step runs of the one-step graph's key (call after ensure_dgraph);
// the segment buffer must hold k + 1 slots
int ensure_kgraphs(lj_ctx* c, double dt, int k)
{
  AdjState* a = c->adj;
  if (a->gexec_k && a->gexec_seg && a->g_k == k) return LJ_OK;
  adj_drop_kgraphs(a);
  LJ_HIP(hipStreamSynchronize(c->stream));
  int rc = capture_dsteps(c, dt, k, false, &a->graph_k, &a->gexec_k);
  if (rc == LJ_OK) rc = capture_dsteps(c, dt, k, true, &a->graph_seg, &a->gexec_seg);
  if (rc != LJ_OK) {
    adj_drop_kgraphs(a);
    return rc;
  }
  a->g_k = k;
  return LJ_OK;
}

int check_traj_args(lj_ctx* c, double dt, int nsteps)
{
  if (nsteps < 1) return fail(LJ_ERR_ARG, "trajectory: nsteps must be >= 1");
  if (!std::isfinite(dt) || dt <= 0.) return fail(LJ_ERR_ARG, "trajectory: dt must be positive");
  if (c->cfg.canonical != LJ_EVN)
    return fail(LJ_ERR_UNSUPPORTED, "trajectory gradients are defined for EVN dynamics only");
  if (c->cfg.boundary == LJ_BC_WALL)
    return fail(LJ_ERR_UNSUPPORTED, "trajectory gradients: hard-wall reflections are not differentiable");
  return LJ_OK;
}

// forward from the context state (converted to FP64): forces at x0, M steps,
// checkpoints every k steps (k <= 0: none); out3 = {L, mean E, mean P}
int forward(lj_ctx* c, double dt, int M, int k, double pt, double wp, double* out3)
{
  AdjState* a = c->adj;
  const int N = c->cfg.N;
  hipStream_t st = c->stream;
  LJ_HIP(launch_dstate_init(st, c->pos, c->vel, N, a->xd, a->vd));
  LJ_TRY(dforces(c, st, a->xd));
  LJ_HIP(launch_dzero4(st, a->acc));
  LJ_TRY(ensure_dgraph(c, dt));
  if (k > 1 && a->seg_cap >= k + 1) LJ_TRY(ensure_kgraphs(c, dt, k));
  if (k > 0) {
    LJ_TRY(dcopy(c, a->ck, a->xd));
    LJ_TRY(dcopy(c, a->ck + N, a->vd));
    LJ_TRY(dcopy(c, a->ck + 2 * N, a->fd));
  }
  // whole checkpoint intervals replay the k-step graph, the rest single steps
  const bool runs = k > 1 && a->gexec_k && a->g_k == k;
  for (int m = 0; m < M;) {
    if (runs && m % k == 0 && m + k <= M) {
      LJ_HIP(hipGraphLaunch(a->gexec_k, st));
      m += k;
    } else {
      LJ_HIP(hipGraphLaunch(a->gexec, st));
      m += 1;
    }
    if (k > 0 && m % k == 0 && m < M) {
      double4* slot = a->ck + (size_t)3 * (m / k) * N;
      LJ_TRY(dcopy(c, slot, a->xd));
      LJ_TRY(dcopy(c, slot + N, a->vd));
      LJ_TRY(dcopy(c, slot + 2 * N, a->fd));
    }
  }
  double acc[4];
  LJ_HIP(hipMemcpyAsync(acc, a->acc, sizeof(acc), hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  const double mE = acc[0] / M, mP = acc[1] / M;
  out3[0] = mE + wp * (mP - pt) * (mP - pt);
  out3[1] = mE;
  out3[2] = mP;
  return LJ_OK;
}

// adjoint pass at FP64 positions x (prep: their internal coordinates are not
// yet in c->q / a->xf) with vector a->z, followed by the reverse sweep's
// per-particle update (`step`, fused into the reduction); the finaliser runs
// when the step's block partials are not accumulated (step.rowacc null)
int adj_pass_d(lj_ctx* c, const double4* x, bool prep, hipStream_t st, const AdjStep& step)
{
  AdjState* a = c->adj;
  if (prep) LJ_HIP(launch_dprep(st, x, c->cfg.N, qscale_of(c), c->q, a->xf));
  LJ_HIP(launch_adj_pass(st, a->plan, c->ps, c->q, a->xf, a->z, a->part));
  const AdjScales sc = make_adj_scales(c->ps);
  LJ_HIP(launch_adj_reduce(st, a->plan, sc, a->part, a->z, a->hz, a->f, a->gw, a->blk, step));
  if (!step.rowacc) LJ_HIP(launch_adj_finalize(st, a->blk, a->nblk, sc, a->scal, step));
  return LJ_OK;
}

}  // namespace

namespace lj {
void adj_free(lj_ctx* c)
{
  AdjState* a = c->adj;
  if (!a) return;
  adj_drop_graph(a);
  dev_free(a->part);
  dev_free(a->z);
  dev_free(a->hz);
  dev_free(a->f);
  dev_free(a->gw);
  dev_free(a->u);
  dev_free(a->lv);
  dev_free(a->dx0);
  dev_free(a->blk);
  dev_free(a->rowacc);
  dev_free(a->scal);
  dev_free(a->theta);
  dev_free(a->acc);
  dev_free(a->xd);
  dev_free(a->vd);
  dev_free(a->fd);
  dev_free(a->xf);
  dev_free(a->ck);
  dev_free(a->seg);
  delete a;
  c->adj = nullptr;
}
}  // namespace lj

extern "C" {

int lj_set_lj_params(lj_ctx* c, double eps, double sigma)
{
  LJ_TRY(ctx_setdev(c));
  if (!(eps > 0.) || !(sigma > 0.)) return fail(LJ_ERR_ARG, "lj_set_lj_params: eps and sigma must be positive");
  // replays enqueued by lj_integrate_async may still run the graphs dropped here
  LJ_HIP(hipStreamSynchronize(c->stream));
  c->cfg.eps = eps;
  c->cfg.sigma = sigma;
  refresh_scalars(c);
  drop_graph(c);
  return LJ_OK;
}

int lj_set_array_device(lj_ctx* c, int type, const float* d_src4, void* stream)
{
  LJ_TRY(ctx_setdev(c));
  float4* d = type == LJ_ARRAY_POS ? c->pos : type == LJ_ARRAY_VEL ? c->vel : type == LJ_ARRAY_FORCE ? c->force : nullptr;
  if (!d || !d_src4) return fail(LJ_ERR_ARG, "lj_set_array_device: bad type or pointer");
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  LJ_HIP(hipMemcpyAsync(d, d_src4, sizeof(float4) * c->cfg.N, hipMemcpyDeviceToDevice, st));
  // an FP64 context steps its double state: widen the new values into it, as lj_set_array does
  if (is_fp64(c)) LJ_HIP(launch_widen(st, d, state64(c, type), c->cfg.N));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

int lj_get_array_device(lj_ctx* c, int type, float* d_dst4, void* stream)
{
  LJ_TRY(ctx_setdev(c));
  float4* d = type == LJ_ARRAY_POS ? c->pos : type == LJ_ARRAY_VEL ? c->vel : type == LJ_ARRAY_FORCE ? c->force : nullptr;
  if (!d || !d_dst4) return fail(LJ_ERR_ARG, "lj_get_array_device: bad type or pointer");
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  LJ_HIP(hipMemcpyAsync(d_dst4, d, sizeof(float4) * c->cfg.N, hipMemcpyDeviceToDevice, st));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

int lj_adjoint_pass_device(lj_ctx* c, const float* d_pos4, const float* d_z4, double* d_hz4,
                           double* d_f4, double* d_gw4, double* d_scal8, void* stream)
{
  LJ_TRY(ctx_setdev(c));
  if (!d_pos4 || !d_z4) return fail(LJ_ERR_ARG, "lj_adjoint_pass_device: null input");
  LJ_TRY(adj_ensure(c, 0, 0));
  AdjState* a = c->adj;
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  const float4* p = reinterpret_cast<const float4*>(d_pos4);
  const float4* z = reinterpret_cast<const float4*>(d_z4);
  if (c->ps.periodic) LJ_HIP(launch_prep_q(st, p, c->q, c->cfg.N, c->ps.qscale));
  LJ_HIP(launch_adj_pass(st, a->plan, c->ps, c->q, p, z, a->part));
  const AdjScales sc = make_adj_scales(c->ps);
  // the plain pass (no fused update) only reads z
  LJ_HIP(launch_adj_reduce(st, a->plan, sc, a->part, const_cast<float4*>(z), a->hz, a->f, a->gw,
                           a->blk));
  LJ_HIP(launch_adj_finalize(st, a->blk, a->nblk, sc, a->scal));
  const size_t b4 = sizeof(double4) * c->cfg.N;
  if (d_hz4) LJ_HIP(hipMemcpyAsync(d_hz4, a->hz, b4, hipMemcpyDeviceToDevice, st));
  if (d_f4) LJ_HIP(hipMemcpyAsync(d_f4, a->f, b4, hipMemcpyDeviceToDevice, st));
  if (d_gw4) LJ_HIP(hipMemcpyAsync(d_gw4, a->gw, b4, hipMemcpyDeviceToDevice, st));
  if (d_scal8) LJ_HIP(hipMemcpyAsync(d_scal8, a->scal, sizeof(double) * kAdjScalars, hipMemcpyDeviceToDevice, st));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

int lj_trajectory_loss(lj_ctx* c, double dt, int nsteps, double p_target, double w_p, double* out3)
{
  LJ_TRY(ctx_setdev(c));
  if (!out3) return fail(LJ_ERR_ARG, "lj_trajectory_loss: null output");
  LJ_TRY(check_traj_args(c, dt, nsteps));
  LJ_TRY(adj_ensure(c, 0, 0));
  return forward(c, dt, nsteps, 0, p_target, w_p, out3);
}

int lj_trajectory_grad(lj_ctx* c, double dt, int nsteps, int checkpoint_every, double p_target,
                       double w_p, double* out3, double* grad_eps_sigma, float* dx0_host4,
                       float* d_dx0_4, double* times_ms3)
{
  LJ_TRY(ctx_setdev(c));
  if (!out3 || !grad_eps_sigma) return fail(LJ_ERR_ARG, "lj_trajectory_grad: null output");
  LJ_TRY(check_traj_args(c, dt, nsteps));
  const int M = nsteps;
  const int k = checkpoint_every <= 0 ? M : std::min(checkpoint_every, M);
  const int ncheck = (M + k - 1) / k;
  LJ_TRY(adj_ensure(c, ncheck, k + 1));
  AdjState* a = c->adj;
  const int N = c->cfg.N;
  hipStream_t st = c->stream;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  for (auto& e : ev) LJ_HIP(hipEventCreate(&e));
  int rc = LJ_OK;
  auto check = [&](hipError_t e) {
    if (e != hipSuccess && rc == LJ_OK)
      rc = fail(LJ_ERR_DEVICE, "lj_trajectory_grad: %s", hipGetErrorString(e));
    return rc == LJ_OK;
  };
  float recompute_ms = 0.f;
  if (check(hipEventRecord(ev[0], st)) && (rc = forward(c, dt, M, k, p_target, w_p, out3)) == LJ_OK &&
      check(hipEventRecord(ev[1], st))) {
    const double vol = std::pow(c->L, 3.);
    const double cP = 2. * w_p * (out3[2] - p_target);
    const double invM = 1. / M;
    const double cw = cP / (3. * vol);               // dP/dx = gradW / (3 vol)
    const double gv = invM * cP * (2. / 3.) / vol;  // g_v = (1/M) cP (2/3) v / vol
    check(launch_adj_zero2(st, a->theta));
    // the block partials of the states m >= 1 add up in rowacc; one finaliser
    // turns their total into the eps / sigma terms at the end of the sweep
    check(launch_fill_u64(st, reinterpret_cast<unsigned long long*>(a->rowacc),
                          (int)((size_t)a->nblk * 8)));
    bool first = true;
    bool prepped = false;  // the next pass's coordinates prepared by the last reduction
    for (int cseg = ncheck - 1; cseg >= 0 && rc == LJ_OK; --cseg) {
      const int s0 = cseg * k, s1 = std::min(M, s0 + k);
      // recompute the segment's states s0..s1 from its checkpoint
      if (!check(hipEventRecord(ev[2], st))) break;
      const double4* ckp = a->ck + (size_t)3 * cseg * N;
      if ((rc = dcopy(c, a->xd, ckp)) || (rc = dcopy(c, a->vd, ckp + N)) || (rc = dcopy(c, a->fd, ckp + 2 * N)))
        break;
      if ((rc = dcopy(c, a->seg, a->xd)) || (rc = dcopy(c, a->seg + N, a->vd))) break;
      if (s1 - s0 == k && a->gexec_seg && a->g_k == k) {
        check(hipGraphLaunch(a->gexec_seg, st));  // k steps, each state into its slot
      } else {
        for (int m = s0 + 1; m <= s1 && rc == LJ_OK; ++m) {
          if (!check(hipGraphLaunch(a->gexec, st))) break;
          double4* slot = a->seg + (size_t)2 * (m - s0) * N;
          if ((rc = dcopy(c, slot, a->xd)) || (rc = dcopy(c, slot + N, a->vd))) break;
        }
      }
      if (rc || !check(hipEventRecord(ev[3], st)) || !check(hipEventSynchronize(ev[3]))) break;
      float ms = 0.f;
      if (!check(hipEventElapsedTime(&ms, ev[2], ev[3]))) break;
      recompute_ms += ms;
      // reverse over m = s1 .. s0+1 (and m = 0 in the first segment)
      const int mlow = cseg == 0 ? 0 : s0 + 1;
      for (int m = s1; m >= mlow && rc == LJ_OK; --m) {
        const double4* xm = a->seg + (size_t)2 * (m - s0) * N;
        if (first) {
          if (!check(launch_adj_init(st, xm + N, N, gv, dt, a->u, a->lv, a->z))) break;
          first = false;
        }
        // the pass, then (fused into its reduction and finaliser) the update of
        // u, lambda and z for state m - 1 (or dL/dx0 at m = 0) and the eps /
        // sigma terms
        AdjStep step;
        step.mode = m >= 1 ? 1 : 2;
        step.invM = invM;
        step.cw = cw;
        step.gv = gv;
        step.dt = dt;
        step.last = m - 1 == 0;
        step.vel_prev = (m - 1 >= 1) ? a->seg + (size_t)2 * (m - 1 - s0) * N + N : nullptr;
        step.u = a->u;
        step.lv = a->lv;
        step.dx0 = a->dx0;
        step.eps = c->cfg.eps;
        step.direct = m >= 1;
        step.g = a->theta;
        step.rowacc = m >= 1 ? a->rowacc : nullptr;
        // the next state of this segment (the next segment's are recomputed
        // first, which overwrites the coordinates)
        if (m - 1 >= mlow) {
          step.next_x = a->seg + (size_t)2 * (m - 1 - s0) * N;
          step.qs = qscale_of(c);
          step.q = c->q;
          step.xf = a->xf;
        }
        if ((rc = adj_pass_d(c, xm, !prepped, st, step)) != LJ_OK) break;
        prepped = step.next_x != nullptr;
      }
      prepped = false;
    }
    if (rc == LJ_OK) {
      AdjStep fin;
      fin.eps = c->cfg.eps;
      fin.invM = invM;
      fin.cw = cw;
      fin.direct = 1;
      fin.g = a->theta;
      check(launch_adj_finalize(st, a->rowacc, a->nblk, make_adj_scales(c->ps), a->scal, fin));
    }
  }
  if (rc == LJ_OK && check(hipEventRecord(ev[3], st)) &&
      check(hipMemcpyAsync(grad_eps_sigma, a->theta, 2 * sizeof(double), hipMemcpyDeviceToHost, st))) {
    if (dx0_host4 || d_dx0_4) {
      std::vector<double4> h(N);
      std::vector<float> hf((size_t)4 * N);
      if (check(hipMemcpyAsync(h.data(), a->dx0, sizeof(double4) * N, hipMemcpyDeviceToHost, st)) &&
          check(hipStreamSynchronize(st))) {
        for (int i = 0; i < N; ++i) {
          hf[4 * i] = (float)h[i].x;
          hf[4 * i + 1] = (float)h[i].y;
          hf[4 * i + 2] = (float)h[i].z;
          hf[4 * i + 3] = 0.f;
        }
        if (dx0_host4) std::memcpy(dx0_host4, hf.data(), sizeof(float) * 4 * N);
        if (d_dx0_4)
          check(hipMemcpyAsync(d_dx0_4, hf.data(), sizeof(float) * 4 * N, hipMemcpyHostToDevice, st));
      }
    }
    if (check(hipStreamSynchronize(st)) && times_ms3) {
      float f_ms = 0.f, all_ms = 0.f;
      check(hipEventElapsedTime(&f_ms, ev[0], ev[1]));
      check(hipEventElapsedTime(&all_ms, ev[1], ev[3]));
      times_ms3[0] = f_ms;
      times_ms3[1] = recompute_ms;
      times_ms3[2] = all_ms - recompute_ms;
    }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

}  // extern "C"
