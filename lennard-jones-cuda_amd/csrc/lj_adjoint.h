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

hipError_t launch_adj_pass(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                           const uint4* q, const float4* posf, const float4* z, float4* part);
hipError_t launch_adj_reduce(hipStream_t st, const ForcePlan& pl, const AdjScales& sc,
                             const float4* part, const float4* z, double4* hz, double4* f,
                             double4* gw, double* blk);
hipError_t launch_adj_finalize(hipStream_t st, const double* blk, int nblk, const AdjScales& sc,
                               double* out);
hipError_t launch_adj_init(hipStream_t st, const double4* vel, int N, double gv, double dt,
                           double4* u, double4* lv, float4* z);
hipError_t launch_adj_update(hipStream_t st, int N, double invM, double cw, double gv, double dt,
                             int last, const double4* hz, const double4* f, const double4* gw,
                             const double4* vel_prev, double4* u, double4* lv, float4* z);
hipError_t launch_adj_final(hipStream_t st, int N, const double4* hz, const double4* u,
                            double4* dx0);
hipError_t launch_adj_theta(hipStream_t st, const double* sc, double eps, double invM, double cw3,
                            int direct, double* g);
hipError_t launch_adj_zero2(hipStream_t st, double* g);

// FP64-state trajectory of the differentiable path
hipError_t launch_dstate_init(hipStream_t st, const float4* pos, const float4* vel, int N,
                              double4* xd, double4* vd);
hipError_t launch_dprep(hipStream_t st, const double4* xd, int N, double qs, uint4* q, float4* xf);
hipError_t launch_ddrift(hipStream_t st, double4* xd, double4* vd, const double4* fd, int N,
                         double dt, double qs, uint4* q, float4* xf);
hipError_t launch_dkick(hipStream_t st, const ForcePlan& pl, double FS, double dt,
                        const float4* fpart, const float* apart, double4* fd, double4* vd,
                        double* blk);
hipError_t launch_dfinal(hipStream_t st, const double* blk, int nblk, double eps, double vol,
                         int accumulate, double* acc);
hipError_t launch_dzero4(hipStream_t st, double* acc);

}  // namespace lj
