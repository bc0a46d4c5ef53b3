// lj_fluct_api.cpp -- C-ABI of the fluctuation counters (lj_fl_*) and of the
// sampling inside the time step (lj_set_step_sampling) of one context.
#include <algorithm>
#include <cmath>
#include <vector>

#include "lj_ctx.h"

using namespace lj;

namespace {

int fl_ready(lj_ctx* c, const char* who)
{
  if (!c->fl.active()) return fail(LJ_ERR_STATE, "%s: no counters configured (lj_fl_configure)", who);
  return LJ_OK;
}

}  // namespace

extern "C" {

int lj_fl_configure(lj_ctx* c, const lj_fl_ladder* ladders, int n)
{
  LJ_TRY(ctx_setdev(c));
  LJ_HIP(hipStreamSynchronize(c->stream));
  // captured steps that sample hold the old tables
  if (c->smp_fl) drop_graph(c);
  LJ_TRY(fl_configure(c->fl, c->stream, 1, &c->L, ladders, n));
  if (!c->fl.active()) c->smp_fl = 0;
  return LJ_OK;
}

int lj_fl_num_entries(lj_ctx* c, int* entries)
{
  if (!c || !entries) return fail(LJ_ERR_ARG, "lj_fl_num_entries: null argument");
  *entries = c->fl.E;
  return LJ_OK;
}

int lj_fl_values(lj_ctx* c, double* values)
{
  if (!c || !values) return fail(LJ_ERR_ARG, "lj_fl_values: null argument");
  std::copy(c->fl.values.begin(), c->fl.values.end(), values);
  return LJ_OK;
}

int lj_fl_sample(lj_ctx* c)
{
  LJ_TRY(ctx_setdev(c));
  LJ_TRY(fl_ready(c, "lj_fl_sample"));
  LJ_HIP(launch_fl_sample(c->stream, c->fl, c->pos, c->vel, c->cfg.N, nullptr, nullptr, 0));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_fl_counts(lj_ctx* c, int64_t* counts)
{
  LJ_TRY(ctx_setdev(c));
  if (!counts) return fail(LJ_ERR_ARG, "lj_fl_counts: null output");
  LJ_TRY(fl_ready(c, "lj_fl_counts"));
  static_assert(sizeof(long long) == sizeof(int64_t), "count width");
  LJ_HIP(hipMemcpyAsync(counts, c->fl.last, sizeof(int64_t) * c->fl.E, hipMemcpyDeviceToHost,
                        c->stream));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_fl_stats(lj_ctx* c, double* stats)
{
  LJ_TRY(ctx_setdev(c));
  if (!stats) return fail(LJ_ERR_ARG, "lj_fl_stats: null output");
  LJ_TRY(fl_ready(c, "lj_fl_stats"));
  LJ_HIP(hipMemcpyAsync(stats, c->fl.stats, sizeof(double) * kFlStat * c->fl.E,
                        hipMemcpyDeviceToHost, c->stream));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_fl_reset(lj_ctx* c)
{
  LJ_TRY(ctx_setdev(c));
  LJ_TRY(fl_ready(c, "lj_fl_reset"));
  LJ_HIP(launch_fl_reset(c->stream, c->fl));
  LJ_HIP(hipStreamSynchronize(c->stream));
  return LJ_OK;
}

int lj_set_step_sampling(lj_ctx* c, const lj_step_sampling* s)
{
  LJ_TRY(ctx_setdev(c));
  if (!s) return fail(LJ_ERR_ARG, "lj_set_step_sampling: null argument");
  if (s->rdf < LJ_SAMPLE_RDF_OFF || s->rdf > LJ_SAMPLE_RDF_SUM)
    return fail(LJ_ERR_ARG, "lj_set_step_sampling: unknown rdf mode %d", s->rdf);
  if (s->fluctuations && !c->fl.active())
    return fail(LJ_ERR_STATE, "lj_set_step_sampling: configure the counters first (lj_fl_configure)");
  hipStream_t st = c->stream;
  LJ_HIP(hipStreamSynchronize(st));
  RdfArgs a{};
  if (s->rdf != LJ_SAMPLE_RDF_OFF) {
    if (is_fp64(c))
      return fail(LJ_ERR_UNSUPPORTED, "lj_set_step_sampling: the step RDF runs in the FP32 modes");
    const int rule = s->rdf_rule, nbins = s->rdf_nbins;
    if (rule < LJ_RDF_REF_CPU || rule > LJ_RDF_RBINS)
      return fail(LJ_ERR_ARG, "lj_set_step_sampling: unknown rule %d", rule);
    if (nbins <= 0 || nbins > 8000)
      return fail(LJ_ERR_ARG, "lj_set_step_sampling: nbins must be in [1, 8000]");
    const float dr2 = s->rdf_dr2 > 0.f ? s->rdf_dr2 : rdf_dr2_rule(c->cfg.N);
    const double rmax = s->rdf_rmax > 0. ? s->rdf_rmax : c->L / 2.;
    if (nbins > c->rdf_step_cap) {
      dev_free(c->rdf_step);
      c->rdf_step_cap = 0;
      LJ_TRY(dev_alloc(&c->rdf_step, nbins));
      c->rdf_step_cap = nbins;
    }
    if (nbins + 1 > c->step_tab_cap) {
      dev_free(c->step_tab);
      c->step_tab_cap = 0;
      LJ_TRY(dev_alloc(&c->step_tab, nbins + 1));
      c->step_tab_cap = nbins + 1;
    }
    std::vector<float> e(nbins + 1);
    rdf_bin_thresholds(rule, nbins, dr2, rmax, e.data());
    LJ_HIP(hipMemcpyAsync(c->step_tab, e.data(), sizeof(float) * (nbins + 1), hipMemcpyHostToDevice,
                          st));
    LJ_HIP(hipStreamSynchronize(st));
    a.pos = c->pos;
    a.N = c->cfg.N;
    a.periodic = c->cfg.boundary == LJ_BC_PERIODIC;
    a.L = c->L;
    a.rule = rule;
    a.nbins = nbins;
    a.tab = c->step_tab;
    a.est_scale = rule == LJ_RDF_RBINS ? (float)(nbins / rmax) : (float)(1.0 / dr2);
    if (rule == LJ_RDF_REF_GPU)
      gpu_minimage_thresholds(c->L, a.thr);
    else
      cpu_minimage_thresholds(c->L, a.thr);
    a.end = rule == LJ_RDF_REF_GPU ? e[nbins - 1] : e[nbins];
    LJ_TRY(attach_rdf_cells(c, a));
    a.sampling = 1;
    a.hist = c->rdf_step;
    LJ_HIP(launch_fill_u64(st, c->rdf_step, nbins));
    LJ_HIP(hipStreamSynchronize(st));
  }
  drop_graph(c);  // the step graphs are captured again with the new sampling
  c->smp_rdf = s->rdf;
  c->smp_args = a;
  c->smp_fl = s->fluctuations ? 1 : 0;
  c->rdf_step_n = 0;
  return LJ_OK;
}

int lj_step_rdf_counts(lj_ctx* c, uint64_t* counts, int64_t* nsamples, int clear)
{
  LJ_TRY(ctx_setdev(c));
  if (c->smp_rdf == LJ_SAMPLE_RDF_OFF)
    return fail(LJ_ERR_STATE, "lj_step_rdf_counts: no RDF sampling set (lj_set_step_sampling)");
  hipStream_t st = c->stream;
  if (counts)
    LJ_HIP(hipMemcpyAsync(counts, c->rdf_step, sizeof(uint64_t) * c->smp_args.nbins,
                          hipMemcpyDeviceToHost, st));
  unsigned bad = 0u;
  if (rdf_uses_cells(c->smp_args))
    LJ_HIP(hipMemcpyAsync(&bad, c->smp_args.cells->result_step, sizeof(unsigned),
                          hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  if (bad) {
    LJ_HIP(hipMemsetAsync(c->smp_args.cells->result_step, 0, sizeof(unsigned), st));
    LJ_HIP(hipStreamSynchronize(st));
    return fail(LJ_ERR_STATE, "lj_step_rdf_counts: particle coordinates left [-L, 2L) during the "
                              "sampled steps; the step RDF is incomplete");
  }
  if (nsamples) *nsamples = c->rdf_step_n;
  if (clear) {
    LJ_HIP(launch_fill_u64(st, c->rdf_step, c->smp_args.nbins));
    LJ_HIP(hipStreamSynchronize(st));
    c->rdf_step_n = 0;
  }
  return LJ_OK;
}

}  // extern "C"
