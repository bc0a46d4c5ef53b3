// lj_internal.h -- shared declarations of the HIP kernels' launchers and the
// host helpers of libljhip.so (not part of the public C-ABI).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lj_hip.h"

namespace lj {

// Observables block kept on the device (same layout as lj_obs).
static_assert(sizeof(lj_obs) == 13 * 8, "lj_obs layout");

// Block-partial fields written by the per-particle kernels; one row of
// kBlkFields doubles per block, reduced in block order by the finalisers.
enum BlkField {
  BLK_K2 = 0,   // sum_i fl32(vx^2+vy^2+vz^2)          (2 K)
  BLK_E = 1,    // sum_i e_i, e_i = sum_j (s/r)^12-(s/r)^6
  BLK_B = 2,    // sum_i b_i, b_i = sum_j (s/r)^6 (2 (s/r)^6 - 1)
  BLK_TV2 = 3,  // TVN: sum_i fl32(|v'|^2)
  BLK_NSUB = 4, // batched boxes: particles inside the sub-volume
  kBlkFields = 5
};

constexpr int kBlock = 256;  // threads per workgroup of the per-particle kernels

// Geometry of one force evaluation (i-tiles x j-slices).
struct ForcePlan {
  int N = 0;
  int S = 1;        // j slices
  int jchunk = 0;   // j per slice
  int Npad = 0;     // row pitch of the partial buffers
  int ipl = 1;      // i per lane
  int gx = 0;       // workgroups along i
  // i rows [i0, iend()): all N (i1 = 0), or one shard's (make_shard_plan)
  int i0 = 0, i1 = 0;
  int iend() const { return i1 > 0 ? i1 : N; }
  bool sharded() const { return i0 != 0 || iend() != N; }
  // pair-once kernel (lj_force_n3.hip): S = ng groups, one workgroup per item
  int n3 = 0;
  int ng = 0;
  int nitems = 0;
  const int2* items = nullptr;  // device table of (g, h) group pairs
  // the time step's variant: nitems_step items (the diagonal items in pairs
  // with W = 2 unless split)
  int nitems_step = 0;
  const int2* items_step = nullptr;
  // workgroups per item: 2 splits every item's j sub-tiles between two
  // workgroups (sizes whose items fill under half the resident workgroups);
  // S = split x ng slots
  int split = 1;
  // 1: the compensated mode's pair-once pass (k_force_n3c: W = 2, not split,
  // diagonal items unpaired)
  int comp = 0;
  // workgroups of the time step's pass, one (E, W) pair of totals each
  int step_wgs() const { return nitems_step * split; }
};
ForcePlan make_force_plan(int N, int jslices_override, int ipl);
// one GPU's share of a force pass sharded by i over nshards <= ceil(N / 64)
// GPUs: rows [64 floor(shard T / nshards), 64 floor((shard + 1) T / nshards))
// (capped at N), T = ceil(N / 64) tiles of 64, against all N on the sliced
// grid, jslices j slices (0: enough to fill the device at that share)
ForcePlan make_shard_plan(int N, int shard, int nshards, int jslices);
ForcePlan make_n3_plan(int N, bool comp = false);
// false when LJ_PAIR_ONCE_COMP=0 keeps the compensated mode on the sliced kernel
bool n3_comp_enabled();
// sizes at which the pair-once kernel serves the time step (N >= 3072)
bool n3_eligible(int N);
// the ng (ng + 1) / 2 group pairs of the pair-once kernel, in launch order
void n3_items(int ng, int2* out);
// the pair-once workgroups of the time step's variant (nitems_step entries)
void n3_step_items(const ForcePlan& pl, int2* out);

// Scalars of the pair kernel for one box.
struct PairScalars {
  int periodic = 1;   // fixed-point minimum image
  double L = 1.0;
  double qscale = 0.; // 2^32 / L (periodic)
  float C1 = 1.f;     // (s/r)^2 = C1 * rcp(r2_internal)
  double FS = 1.;     // force = FS * sum s_raw * d_raw
  float thr = 0.f;    // guard-all threshold in internal r^2 units (0 = self by index only)
  double eps = 1., sigma = 1.;
};
PairScalars make_pair_scalars(double L, int periodic, double eps, double sigma, bool guard_all);

// Kernel launchers (enqueue only).
hipError_t launch_prep_q(hipStream_t st, const float4* pos, uint4* q, int N, double qscale);
// ties: resolve exact half-box separations as the reference does (single
// evaluations of arbitrary states); the step graph uses ties = false
// step_ew (pair-once plans only, may be null): the time step's variant, which
// writes force-only partials and per-workgroup (E, W) totals to step_ew
struct StepArgs;
// the time step's pair-once pass on its items [first, first + count) (W = 2,
// paired diagonal items, not split): one GPU's share when items are dealt out
hipError_t launch_force_n3_step_items(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                                      const uint4* q, const float4* posf, float4* fpart, double* ew,
                                      int first, int count);
hipError_t launch_force(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                        const uint4* q, const float4* posf, float4* fpart, float* apart,
                        bool compensated, bool ties, double* step_ew = nullptr);
hipError_t launch_force_n3(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                           const uint4* q, const float4* posf, float4* fpart, float* apart,
                           bool ties, double* ew);
// sum slices -> force (w = e_i); block partials of e, b
hipError_t launch_reduce_forces(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                                const float4* fpart, const float* apart, float4* force,
                                double* blk);
// sum the rows of blk (nblk x kBlkFields) -> ev = {E_pot, W}; pressure_f (may be
// null) = (float)(4 eps sum b)  [reference GPU 'pressure' convention]
hipError_t launch_finalize_forces(hipStream_t st, const double* blk, int nblk, double eps,
                                  double* ev, float* pressure_f);

struct StepArgs {
  float4* pos;
  float4* vel;
  float4* force;
  float4* tF;
  uint4* q;
  const float4* fpart;
  const float* apart;
  double* blk;
  lj_obs* obs;
  double* hist;      // 5 per step
  int64_t* hist_base;
  int hist_cap;
  int N;
  double dt, L, T0, rho, eps, qscale, FS;
  int bc, canonical, S, Npad;
  // nitems > 0: the force pass wrote force-only partials and per-workgroup
  // totals ew[2 k] (sum of e over ordered pairs), ew[2 k + 1] (sum of w)
  const double* ew;
  int nitems;
  // block rows of the step's sums (K2, TV2, E, B): one per kRowParticles
  // particles in the FP32 step kernels, one per kBlock in the FP64 ones
  int nrow;
};
// steps per replay of the multi-step graphs (single box and batch): each graph
// launch costs about 8 us of idle device time between replays, against none
// between the kernels inside one graph.  Long single-box runs replay a graph
// of kLongGraphSteps first (N = 16384, TVN: 80.8 against 82.1 us per step).
constexpr int kGraphSteps = 8;
constexpr int kLongGraphSteps = 32;
// particles per block row of the FP32 time-step kernels
constexpr int kRowParticles = 64;
inline int step_rows(int N) { return (N + kRowParticles - 1) / kRowParticles; }
hipError_t launch_drift(hipStream_t st, const StepArgs& a);
// drift_next: also the next step's drift (as launch_drift) for each particle
// fin != null (pair-once time step only): one extra block computes the
// observables of the previous step of the run
hipError_t launch_kick_evn(hipStream_t st, const StepArgs& a, bool drift_next = false,
                           const StepArgs* fin = nullptr);
hipError_t launch_tvn_a(hipStream_t st, const StepArgs& a, const StepArgs* fin = nullptr);
hipError_t launch_tvn_b(hipStream_t st, const StepArgs& a, bool drift_next = false);
// observables of the step; advance t/step, write history
hipError_t launch_finalize_step(hipStream_t st, const StepArgs& a);
// observables of the current state (no time advance): K from vel, V/W from blk
hipError_t launch_params(hipStream_t st, const StepArgs& a, int include_force_terms);
hipError_t launch_set_hist_base(hipStream_t st, lj_obs* obs, int64_t* base);
// velocity rescale by sqrt(T0/Tkin) or by a host factor (factor > 0)
hipError_t launch_kinetic_partials(hipStream_t st, const float4* vel, int N, double* blk);
hipError_t launch_scale_velocities(hipStream_t st, float4* vel, int N, const double* blk,
                                   int nblk, double T0, double factor, lj_obs* obs);
hipError_t launch_apply_bc(hipStream_t st, float4* pos, float4* vel, int N, int bc, double L);
hipError_t launch_reset_averaging(hipStream_t st, lj_obs* obs);
// obs->V, obs->W from ev = {E_pot, W}
hipError_t launch_store_vw(hipStream_t st, const double* ev, lj_obs* obs);
// obs->t = 0, obs->step = 0
hipError_t launch_reset_clock(hipStream_t st, lj_obs* obs);

// Start and stop events for the force-pass launches of this thread (both null
// except inside lj_time_force_kernel_stamped): the launchers hand them to
// hipExtLaunchKernel, which stamps them with the kernel's own start and end
struct LaunchStamp {
  hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchStamp tl_force_stamp;

// RDF
// Device workspace of the cell-list RDF pass (lj_rdf_cells.hip), owned by a
// context and grown by rdf_cells_ensure before any capture
constexpr int kRdfReps = 64;       // replicas of the workgroup histograms
constexpr int kRdfMaxCells = 64;   // cells per axis at most
constexpr int kCellScanMax = 1 << 16;  // (cells x blocks) of the LDS-counted sort
struct RdfCellsWs {
  int* count = nullptr;            // [cap_cells]: particles per cell (cleared by the scan or
                                   // the reduction)
  int* start = nullptr;            // [cap_cells + 1]
  int* cellof = nullptr;           // [cap_n]
  int* rank = nullptr;             // [cap_n]
  uint4* sorted = nullptr;         // [cap_n]: (q.x, q.y, q.z, particle)
  unsigned* reps = nullptr;        // [kRdfReps][cap_bins]
  unsigned long long* other = nullptr;  // [kRdfReps]
  unsigned* flag = nullptr;        // coordinates beyond [-L, 2 L) in the current pass
  unsigned* result = nullptr;      // a flagged lj_rdf pass, for the host
  unsigned* result_step = nullptr; // a flagged pass of the per-step sampling, for the host
  int* bcnt = nullptr;             // [cap_m]: each block's offset within each cell, block-major
  int cap_n = 0, cap_cells = 0, cap_bins = 0, cap_m = 0;
};
// cells per axis for a periodic range ending at r^2 = end, or 0 (fewer than 3)
int rdf_cells_per_axis(double L, double end);
// cells per axis of the periodic layout the pass takes (0: none), and its
// reach (1 or 2)
int rdf_cells_layout(double L, double end, int* reach);
// LJ_OK or a status; st: the owner's stream (its work is waited for, and the
// new arrays are cleared on it); *grown: whether the arrays were reallocated
int rdf_cells_ensure(RdfCellsWs& w, int N, int ncell, int nbins, hipStream_t st, bool* grown);
void rdf_cells_free(RdfCellsWs& w);

struct RdfArgs {
  const float4* pos;
  int N;
  int periodic;
  double L;
  int rule;
  int nbins;
  const float* tab;  // rdf_bin_thresholds (nbins + 1)
  float est_scale;   // bin estimate: r^2 est_scale (REF rules), sqrt(r^2) est_scale (RBINS)
  float thr[4];      // minimum-image thresholds of the rule
  double end;        // the counted range's end, tab[nbins - 1] (REF_GPU) or tab[nbins]
  unsigned long long* hist;
  const RdfCellsWs* cells;  // may be null: no cell-list pass
  int sampling;             // the per-step sampling's pass (its own result word)
};
hipError_t launch_rdf(hipStream_t st, const RdfArgs& a);
hipError_t launch_rdf_cells(hipStream_t st, const RdfArgs& a);
// cells per axis (0: none) and reach of the cell-list pass for these
// arguments, periodic or hard-wall
int rdf_cells_grid(const RdfArgs& a, int* reach);
// whether launch_rdf takes the cell-list pass for these arguments
bool rdf_uses_cells(const RdfArgs& a);
hipError_t launch_fill_u64(hipStream_t st, unsigned long long* p, int n);
// speed histogram (MDSystem.cpp:651-694); hist is added to (clear it first)
// shear stress of the reference CPU path (MDSystem.cpp:299,309,335,353) over
// the sliced grid of pl; part: pl.gx * pl.S doubles; out: one double
hipError_t launch_shear(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                        const uint4* q, const float4* pos, const float4* vel, double volume,
                        double* part, double* out);
hipError_t launch_velocity_hist(hipStream_t st, const float4* vel, int N, double step, int nbins,
                                unsigned long long* hist);

// ---------------------------------------------------------------------------
// batched ensemble (lj_batch.hip)

// Per-box constants of a batched ensemble (device array, one per box).
struct BoxParam {
  int N, off;        // particles, offset of the box in the flat arrays
  int S, jchunk;     // j slices of the force pass
  int Npad;          // row pitch of the box's partial buffers
  int pbase;         // offset of the box's partial rows (elements)
  int blk0, nblk;    // first block row and number of block rows
  int ew0, nitem;    // pair-once items of the box: their (E, W) totals at ew[2 (ew0 + k)]
  double L, qscale, FS, rho, T0;
  float C1, pad;
  double sub_lo, sub_hi;  // sub-volume cube [lo, hi]^3 of the fluctuation counts
};

// One work item of the batched force pass: box, i-tile start (local), slice;
// with the pair-once decomposition (box, group g, group h).
constexpr int kPairOnceGroup = 256;  // group size of the batched pair-once pass

struct BatchItem {
  int box, i0, s;
  int e;  // pair-once: the item's slot in the per-box runs of (E, W) totals
};

// per-box accumulated fluctuation statistics (LJ_BATCH_NSTAT doubles per box)
enum BatchStat {
  ST_SAMPLES = 0, ST_U, ST_U2, ST_V, ST_V2, ST_K, ST_K2, ST_P, ST_P2,
  ST_N1, ST_N2, ST_N3, ST_N4, kNStat
};

struct BatchArgs {
  float4* pos;
  float4* vel;
  float4* force;
  float4* tF;
  uint4* q;
  float4* fpart;
  float* apart;
  double* blk;
  lj_obs* obs;          // per box
  double* stats;        // per box, kNStat
  const BoxParam* box;
  const int2* pblock;   // per block: (box, local start)
  const BatchItem* items;
  unsigned* counter;    // work-queue head
  int nboxes, nblocks, nitems;
  double dt, eps;
  int bc, canonical;
  int ipl, persistent_grid;
  int pair_once;        // items are (box, g, h) group pairs of 256 (k_batch_force_n3)
  // lean (the time step with pair_once): the force pass writes force-only
  // partials and per-item totals ew[2 e] = sum of e, ew[2 e + 1] = sum of w over
  // ordered pairs; the kick kernels then read no apart, and the finaliser adds
  // each box's item totals
  int lean;
  double* ew;
  // per-step history (may be null): row k of box b at hist[5 (k nboxes + b)]
  // = {U, T, K, V, P} for k = obs[b].step - hist_base[b] < hist_cap
  double* hist;
  const int64_t* hist_base;
  int hist_cap;
};
hipError_t launch_batch_drift(hipStream_t st, const BatchArgs& a);
hipError_t launch_batch_force(hipStream_t st, const BatchArgs& a, bool periodic, bool ties);
hipError_t launch_batch_force_n3(hipStream_t st, const BatchArgs& a, bool periodic, bool ties);
// drift_next: also the next step's drift of each particle (as launch_batch_drift)
hipError_t launch_batch_kick_evn(hipStream_t st, const BatchArgs& a, bool drift_next = false);
hipError_t launch_batch_tvn_a(hipStream_t st, const BatchArgs& a);
hipError_t launch_batch_tvn_b(hipStream_t st, const BatchArgs& a, bool drift_next = false);
hipError_t launch_batch_finalize(hipStream_t st, const BatchArgs& a, int advance);
hipError_t launch_batch_prep_q(hipStream_t st, const BatchArgs& a);
hipError_t launch_batch_reduce_forces(hipStream_t st, const BatchArgs& a);
hipError_t launch_batch_kinetic(hipStream_t st, const BatchArgs& a);
hipError_t launch_batch_zero_stats(hipStream_t st, const BatchArgs& a);
// hist_base[b] = obs[b].step (the next step writes history row 0)
hipError_t launch_batch_set_hist_base(hipStream_t st, const BatchArgs& a, int64_t* hist_base);

// host helpers (lj_host.cpp)
void lattice_positions(int N, double L, float* pos4);
void gaussian_velocities(int N, double T0, uint64_t seed, float* vel4);
// the reference's SampleInitialConditions velocities (Maxwell speeds by
// rejection, isotropic directions), seeded; momentum removed, T rescaled
void maxwell_velocities(int N, double T0, uint64_t seed, float* vel4);
double box_length(int N, double rho);
float rdf_dr2_rule(int N);
// smallest positive float x with fast_round((float)(x / L)) >= m, m = 1..4
void cpu_minimage_thresholds(double L, float thr[4]);
// the same for the GPU rule: thr[m-1] = least x >= 0 with lround(x / L) >= m (float)
void gpu_minimage_thresholds(double L, float thr[4]);
void rdf_edges2(int nbins, double rmax, float* edges2);
// per-bin r^2 thresholds of an RDF rule: T[k] = the least float r^2 whose bin
// is >= k (k = 0..nbins, T[0] = 0); r^2 >= T[nbins] is dropped (REF_CPU,
// RBINS) or, REF_GPU, T[nbins] = +inf (clamped into the last bin)
void rdf_bin_thresholds(int rule, int nbins, float dr2, double rmax, float* T);

}  // namespace lj
