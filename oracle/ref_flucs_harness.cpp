// ref_flucs_harness.cpp -- C entry points over the reference's fluctuation
// counters and time averages.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile into
// oracle/_ref/libref_cpu.so together with the reference sources where they lie:
//   src/tasks/run-fluctuations/include/run-fluctuations-aux.h (GetNSubsystem,
//     GetNsubVz, GetNSubsystemBatch, GetNsubVzBatch, CoordFlucsAverage,
//     MomentumFlucsAverage), which pulls in
//   src/tasks/auxiliary/time-average-aux.h (TimeAverage) and
//   src/extra/sample-moments/NumberStatistics.h (moment statistics).
// The counters read the public h_Pos / h_Vel of an MDSystem made by
// ref_harness.cpp (refh_create / refh_set_state).
#include <cstring>

#include "run-fluctuations-aux.h"

using namespace RunFluctuationsFunctions;

extern "C" {

int refh_fl_nsub_batch(void* h, double alpha_step, int type, int* out, int cap)
{
  std::vector<int> c = GetNSubsystemBatch(*static_cast<MDSystem*>(h), alpha_step, type);
  for (size_t k = 0; k < c.size() && (int)k < cap; ++k) out[k] = c[k];
  return (int)c.size();
}

int refh_fl_vz_batch(void* h, double vcut_max, double alpha_step, int type, int* out, int cap)
{
  std::vector<int> c = GetNsubVzBatch(*static_cast<MDSystem*>(h), vcut_max, alpha_step, type);
  for (size_t k = 0; k < c.size() && (int)k < cap; ++k) out[k] = c[k];
  return (int)c.size();
}

int refh_fl_nsub(void* h, double fraction, int type)
{
  return GetNSubsystem(*static_cast<MDSystem*>(h), fraction, type);
}

int refh_fl_vz(void* h, double vcut, int type)
{
  return GetNsubVz(*static_cast<MDSystem*>(h), vcut, type);
}

// MDSystem::Fluctuations (MDSystem.cpp:708-730) once; out5 = {momN, momN2,
// momN3, momN4, momiters} after the call (it also prints a line to stdout)
void refh_fl_fluctuations(void* h, double fraction, double* out5)
{
  MDSystem* s = static_cast<MDSystem*>(h);
  s->Fluctuations(fraction);
  out5[0] = s->momN;
  out5[1] = s->momN2;
  out5[2] = s->momN3;
  out5[3] = s->momN4;
  out5[4] = s->momiters;
}

// TimeAverage (time-average-aux.h:27-67) over n values; out7 = {mean,
// variance, s, mean error (x sqrt s), scaled variance, scaled variance error,
// plain mean error}
void refh_tavg(const double* vals, int n, double* out7)
{
  TimeAverage a;
  for (int k = 0; k < n; ++k) a.AddObservation(vals[k]);
  out7[0] = a.GetMean();
  out7[1] = a.stats.GetVariance();
  out7[2] = a.GetS();
  out7[3] = a.GetMeanError();
  out7[4] = a.stats.GetScaledVariance();
  out7[5] = a.stats.GetScaledVarianceError();
  out7[6] = a.stats.GetMeanError();
}

}  // extern "C"
