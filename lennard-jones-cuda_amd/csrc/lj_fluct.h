// lj_fluct.h -- fluctuation observables on the device: particle counts of
// sub-volumes and velocity windows and their accumulated moments, for one box
// (lj_ctx) or every box of a batch (lj_batch).
//
// A ladder is one counter of the reference's fluctuation tasks, evaluated with
// its own double arithmetic on the float state:
//   run-fluctuations-aux.h:191-241  GetNSubsystemBatch (cubes, slabs)
//   run-fluctuations-aux.h:246-278  GetNsubVzBatch     (velocity windows)
//   run-fluctuations-aux.h:128-186  GetNSubsystem, GetNsubVz (one fraction)
//   MDSystem.cpp:708-724            MDSystem::Fluctuations (closed cube)
// Each ladder has m entries; entry k counts the particles inside window k
// (cumulative, as the reference returns them).  A sample adds each entry's
// count n to its statistics: sum n^k (k = 0..4, successive products as
// NumberStatistics.h:608-617), the lag-one product sum and the last value
// (TimeAverage, time-average-aux.h:38-48), from which the host forms means,
// variances, scaled variances and their errors.
#pragma once

#include <vector>

#include "lj_internal.h"

namespace lj {

// statistics per entry (doubles)
enum FlStat {
  FL_S0 = 0,  // number of samples
  FL_S1, FL_S2, FL_S3, FL_S4,
  FL_CORR,    // sum over samples t >= 2 of n_t n_{t-1}
  FL_LAST,    // n of the last sample
  FL_ITERS,   // samples (TimeAverage::iters)
  kFlStat
};

// how a ladder turns a particle into an entry index
enum FlKind {
  FL_TABLE_CUBE = 0,  // v = max_a |x_a - L/2| / 0.5; index = #{t_k <= v} (upper_bound)
  FL_TABLE_COORD = 1, // v = x_axis;   index = #{t_k <= v}
  FL_TABLE_VABS = 2,  // v = |v_axis|; index = #{t_k <= v}
  FL_DIV_COORD = 3,   // index = (int)((x_axis / L) / step)
  FL_DIV_VEL = 4,     // index = (int)(|v_axis / vmax| / step)
  FL_CLOSED_CUBE = 5  // index 0 when lo <= x_a <= hi on every axis
};

// one ladder of one box (device table, nbox x nladders)
struct FlLadderDev {
  int kind, axis, m, off;  // off: first entry of the ladder in a box's entries
  double step, scale;      // DIV kinds: step and L (coord) or vmax (vel)
  double lo, hi;           // FL_CLOSED_CUBE bounds
};

struct FlSampler {
  int nbox = 0, nl = 0, E = 0;  // boxes, ladders, entries per box
  int maxm = 0;
  std::vector<lj_fl_ladder> spec;
  std::vector<double> values;   // per entry: its alpha (fraction, or vcut for one window)
  FlLadderDev* lad = nullptr;   // [nbox * nl]
  double* tab = nullptr;        // [nbox * E] thresholds of the table kinds
  unsigned* counts = nullptr;   // [nbox * E] per-sample bins; zero between samples
  long long* last = nullptr;    // [nbox * E] cumulative counts of the last sample
  double* stats = nullptr;      // [nbox * E * kFlStat]
  bool active() const { return nl > 0; }
};

// Builds the ladders for boxes of sides L[0..nbox) (validates spec) and
// allocates and clears the device tables; synchronous on st.
int fl_configure(FlSampler& s, hipStream_t st, int nbox, const double* L, const lj_fl_ladder* spec,
                 int nl);
void fl_free(FlSampler& s);
// zero the statistics (kernel, no memset node)
hipError_t launch_fl_reset(hipStream_t st, const FlSampler& s);
// one sample of the state: count, then accumulate.  box == null: one box of N
// particles; otherwise the batch's boxes (BoxParam off/N) with blocks mapped
// by pblock (nblocks entries of (box, first particle))
hipError_t launch_fl_sample(hipStream_t st, const FlSampler& s, const float4* pos, const float4* vel,
                            int N, const BoxParam* box, const int2* pblock, int nblocks);

}  // namespace lj
