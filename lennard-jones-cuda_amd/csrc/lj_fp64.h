// lj_fp64.h -- launchers of the FP64 precision mode (lj_fp64.hip): double
// state, double pair arithmetic, the reference's integrator formulas in double.
#pragma once

#include "lj_internal.h"

namespace lj {

struct Step64 {
  double4* x;        // FP64 state (authoritative in this mode)
  double4* v;
  double4* f;        // forces of x, (fx, fy, fz, e_i)
  double4* tf;       // TVN half-sum of old and new forces
  float4* pos;       // FP32 mirrors of the state at the boundary
  float4* vel;
  float4* force;
  const double4* part;  // slice partials (fx, fy, fz, b)
  const double* apart;  // slice partials a
  double* blk;
  int N, S, Npad, bc;
  double dt, L, T0, FS;
};

hipError_t launch_force64(hipStream_t st, const ForcePlan& pl, const double4* x, bool periodic,
                          double L, double sigma, double4* part, double* apart);
// slice sums -> f, FP32 force mirror, block partials (E, B)
hipError_t launch_reduce64(hipStream_t st, const Step64& a);
hipError_t launch_drift64(hipStream_t st, const Step64& a, int canonical);
hipError_t launch_kick64(hipStream_t st, const Step64& a);
hipError_t launch_tvn64_a(hipStream_t st, const Step64& a);
hipError_t launch_tvn64_b(hipStream_t st, const Step64& a);
// float4 -> double4 (state upload through the FP32 arrays) and back
hipError_t launch_widen(hipStream_t st, const float4* src, double4* dst, int N);
hipError_t launch_narrow(hipStream_t st, const double4* src, float4* dst, int N);
// block partials of sum v^2 (BLK_K2) from the FP64 velocities
hipError_t launch_kinetic64(hipStream_t st, const double4* v, int N, double* blk);

// RenormalizeVelocities on the double velocities (factor <= 0: sqrt(T0/Tkin)
// from the BLK_K2 partials) and the boundary conditions on the double state;
// both refresh the float mirrors
hipError_t launch_scale64(hipStream_t st, double4* v, float4* vel, int N, const double* blk,
                          int nblk, double T0, double factor, lj_obs* obs);
hipError_t launch_bc64(hipStream_t st, double4* x, double4* v, float4* pos, float4* vel, int N,
                       int bc, double L);

}  // namespace lj
