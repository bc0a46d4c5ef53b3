// lj_adjoint.h -- launchers of the adjoint kernels (lj_adjoint.hip).
#pragma once

#include "lj_internal.h"

namespace lj {

// scalar outputs of one adjoint pair pass (double[8])
enum AdjScalar { ADJ_E = 0, ADJ_W, ADJ_E_SIGMA, ADJ_W_SIGMA, ADJ_FSIGMA_Z, ADJ_F_Z, kAdjScalars = 8 };

struct AdjScales {
  double eps, sigma;
  double h;     // H z scale: 24 eps / s^2
  double f;     // force:      24 eps u / s^2
  double g;     // grad W:     48 eps u / s^2
  double fsig;  // dF/dsigma:  24 eps u / s^3
};
AdjScales make_adj_scales(const PairScalars& ps);
// block rows of the adjoint reduction and FP64 kick kernels (64 particles each)
inline int adj_rows(int N) { return (N + 63) / 64; }

// What the reverse sweep fuses into the reduction and finalising kernels of
// one adjoint pass (mode 0: nothing, the plain pass of lj_adjoint_pass_device):
//   mode 1 (states m >= 1): the per-particle update of u, lambda_v and z;
//   mode 2 (state 0): dL/dx0 = u + H z into dx0;
//   g != null: the eps / sigma terms added to g after the scalars.
struct AdjStep {
  int mode = 0;
  double invM = 0., cw = 0., gv = 0., dt = 0.;
  int last = 0;
  const double4* vel_prev = nullptr;
  double4* u = nullptr;
  double4* lv = nullptr;
  double4* dx0 = nullptr;
  double eps = 1.;
  int direct = 0;
  double* g = nullptr;
  // reduction kernel: the next state's internal coordinates (as k_dprep) from
  // next_x (may be null), and the block partials added to the rows of rowacc
  // instead of written to blk (may be null: written)
  const double4* next_x = nullptr;
  double qs = 0.;
  uint4* q = nullptr;
  float4* xf = nullptr;
  double* rowacc = nullptr;
};

// the pair pass's geometry for N particles (slices of the j range)
ForcePlan adj_force_plan(int N);
hipError_t launch_adj_pass(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                           const uint4* q, const float4* posf, const float4* z, float4* part);
// z: the pass's adjoint vector; with step.mode == 1 the kernel then writes the
// next state's vector into it (each thread its own particle, after reading it)
hipError_t launch_adj_reduce(hipStream_t st, const ForcePlan& pl, const AdjScales& sc,
                             const float4* part, float4* z, double4* hz, double4* f,
                             double4* gw, double* blk, const AdjStep& step = AdjStep());
hipError_t launch_adj_finalize(hipStream_t st, const double* blk, int nblk, const AdjScales& sc,
                               double* out, const AdjStep& step = AdjStep());
hipError_t launch_adj_init(hipStream_t st, const double4* vel, int N, double gv, double dt,
                           double4* u, double4* lv, float4* z);
hipError_t launch_adj_zero2(hipStream_t st, double* g);

// FP64-state trajectory of the differentiable path
hipError_t launch_dstate_init(hipStream_t st, const float4* pos, const float4* vel, int N,
                              double4* xd, double4* vd);
hipError_t launch_dprep(hipStream_t st, const double4* xd, int N, double qs, uint4* q, float4* xf);
hipError_t launch_ddrift(hipStream_t st, double4* xd, double4* vd, const double4* fd, int N,
                         double dt, double qs, uint4* q, float4* xf);
// slot_x / slot_v (may be null): the state after the kick, positions xd and
// the kicked velocities, is also written there (the recompute's segment slots)
hipError_t launch_dkick(hipStream_t st, const ForcePlan& pl, double FS, double dt,
                        const float4* fpart, const float* apart, double4* fd, double4* vd,
                        double* blk, const double4* xd = nullptr, double4* slot_x = nullptr,
                        double4* slot_v = nullptr);
hipError_t launch_dfinal(hipStream_t st, const double* blk, int nblk, double eps, double vol,
                         int accumulate, double* acc);
hipError_t launch_dzero4(hipStream_t st, double* acc);

}  // namespace lj
