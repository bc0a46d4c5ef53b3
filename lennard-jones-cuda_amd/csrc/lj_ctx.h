// lj_ctx.h -- the context object behind lj_ctx* and the internal helpers
// shared by the C-ABI implementation files.
#pragma once

#include "lj_fluct.h"
#include "lj_internal.h"
#include "lj_util.h"

struct AdjState;  // adjoint scratch (lj_adjoint_api.cpp), allocated on first use

struct lj_ctx {
  lj_config cfg{};
  double L = 0.;
  hipStream_t stream = nullptr;
  lj::ForcePlan plan;   // sliced force pass (also the adjoint pass geometry)
  lj::ForcePlan plan3;  // pair-once force pass, when enabled (FP32 mode, large N)
  int2* n3items = nullptr;
  double* step_ew = nullptr;  // per-workgroup (E, W) of the step's pair-once pass
  lj::PairScalars ps;
  float4* pos = nullptr;
  float4* vel = nullptr;
  float4* force = nullptr;
  float4* tF = nullptr;
  uint4* q = nullptr;
  float4* fpart = nullptr;
  float* apart = nullptr;
  double* blk = nullptr;
  int nblk = 0;
  lj_obs* obs = nullptr;
  double* hist = nullptr;
  int hist_cap = 0;
  int64_t* hist_base = nullptr;
  double* ev = nullptr;
  unsigned long long* rdf = nullptr;
  int rdf_cap = 0;
  int rdf_nbins = 0, rdf_rule = -1;
  float rdf_dr2 = 0.f;
  double rdf_rmax = 0.;
  float* edges2 = nullptr;
  int edges_cap = 0;
  unsigned long long* vh = nullptr;  // speed histogram counts
  int vh_cap = 0;
  double* shear = nullptr;  // shear-stress workgroup parts, then the result
  int shear_cap = 0;
  // one shard's force pass (lj_forces_shard): partial entries, block rows of
  // its reduction and its two totals
  float4* sh_fpart = nullptr;
  float* sh_apart = nullptr;
  double* sh_blk = nullptr;
  double* sh_ev = nullptr;
  size_t sh_cap = 0;
  int sh_blk_cap = 0;
  // > 1: the step's pair-once pass runs as that many consecutive launches on
  // contiguous item ranges, as that many GPUs would run their shares
  // (LJ_EMULATE_SHARDS, read at creation; a check of the sharded design)
  int emulate_shards = 1;
  hipGraph_t graph = nullptr;        // one step
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph_run = nullptr;    // kGraphSteps steps back to back
  hipGraphExec_t gexec_run = nullptr;
  hipGraph_t graph_long = nullptr;   // kLongGraphSteps steps back to back
  hipGraphExec_t gexec_long = nullptr;
  double g_dt = 0.;
  int g_canon = -1, g_bc = -1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  AdjState* adj = nullptr;
  // LJ_PREC_FP64: double state (authoritative; pos/vel/force are its mirrors)
  double4* xd = nullptr;
  double4* vd = nullptr;
  double4* fd = nullptr;
  double4* tfd = nullptr;
  double4* part64 = nullptr;
  double* apart64 = nullptr;
  // fluctuation counters (lj_fl_*)
  lj::FlSampler fl;
  // sampling inside the step graphs (lj_set_step_sampling)
  int smp_rdf = LJ_SAMPLE_RDF_OFF;
  int smp_fl = 0;
  lj::RdfArgs smp_args{};              // RDF pass of each step (hist = rdf_step)
  unsigned long long* rdf_step = nullptr;
  int rdf_step_cap = 0;
  float* step_tab = nullptr;           // bin thresholds of the step's RDF pass
  int step_tab_cap = 0;
  int64_t rdf_step_n = 0;              // steps the step histogram holds
  lj::RdfCellsWs rdf_cells;            // workspace of the cell-list RDF pass
  bool sampling() const { return smp_rdf != LJ_SAMPLE_RDF_OFF || smp_fl != 0; }
};

namespace lj {
int ctx_setdev(lj_ctx* c);
// step graphs (one step, kGraphSteps steps) captured once per (dt, ensemble,
// boundary) and replayed
int ensure_graph(lj_ctx* c, double dt);
// enqueues nsteps steps through the captured graphs
int launch_steps(lj_ctx* c, int nsteps);
// device forces, E and W of the context positions into c->force and c->obs
int compute_forces_internal(lj_ctx* c);
void refresh_scalars(lj_ctx* c);
void drop_graph(lj_ctx* c);
// releases the adjoint scratch of a context
void adj_free(lj_ctx* c);
inline bool is_fp64(const lj_ctx* c) { return c->cfg.precision == LJ_PREC_FP64; }
// the double state array of LJ_ARRAY_POS / VEL / FORCE (LJ_PREC_FP64 contexts)
double4* state64(lj_ctx* c, int type);
// plan of the force pass that fills fpart/apart
inline const ForcePlan& fplan(const lj_ctx* c) { return c->plan3.n3 ? c->plan3 : c->plan; }
// gives an RDF pass of the context the cell-list workspace where that pass
// applies (at most 1024 bins; periodic boxes with 3 or more cells per axis,
// hard-wall boxes with 6 or more), grown first
inline int attach_rdf_cells(lj_ctx* c, RdfArgs& a)
{
  a.cells = nullptr;
  if (a.nbins > 1024) return LJ_OK;
  int reach = 1;
  const int nc = rdf_cells_grid(a, &reach);
  if (nc <= 0) return LJ_OK;
  bool grown = false;
  const int rc = rdf_cells_ensure(c->rdf_cells, a.N, nc * nc * nc, a.nbins, c->stream, &grown);
  if (rc != LJ_OK) return rc;
  // reallocated: captured steps that sample the RDF hold the old arrays'
  // addresses (lj_rdf and the per-step sampling share the workspace); they
  // are captured again on the next run
  if (grown) drop_graph(c);
  a.cells = &c->rdf_cells;
  return LJ_OK;
}
}  // namespace lj
