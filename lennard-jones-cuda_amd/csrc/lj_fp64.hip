// lj_fp64.hip -- the FP64 precision mode: double positions, velocities and
// forces, double pair arithmetic, and the reference integrator's formulas
// (MDSystem.cpp:442-510) evaluated in double. It is the device reference of
// the precision study: the FP32 and compensated modes are measured against it
// over long trajectories that a CPU cannot follow at N = 131072.
//
// Pair pass: the single-box decomposition (one i per lane, grid = i-tiles x
// j-slices, j streamed through SGPRs by uniform scalar loads); minimum image
// dx -= L rint(dx / L) in double; self pair excluded by index (MDSystem.cpp:267).
// Slice partials are summed per particle in fixed order.
#include "lj_device.h"
#include "lj_fp64.h"

namespace lj {

namespace {

using namespace dev;

// 1 / x in double: the hardware estimate (v_rcp_f64) refined by two Newton
// steps, within an ulp or two of the quotient (the IEEE division's scaling
// and fix-up sequence cost several times as much in the pair loop)
__device__ __forceinline__ double rcp64(double x)
{
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.);
  return __builtin_fma(y, e, y);
}

// one pair (i, j) into the lane's sums; GUARD: j may be i (the self pair,
// excluded)
template <bool PERIODIC, bool GUARD>
__device__ __forceinline__ void pair64(const double4& xi, const double4& xj, int i, int j, double L,
                                       double invL, double s2, double& fx, double& fy, double& fz,
                                       double& a, double& b)
{
  double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
  if constexpr (PERIODIC) {
    dx -= L * __builtin_rint(dx * invL);
    dy -= L * __builtin_rint(dy * invL);
    dz -= L * __builtin_rint(dz * invL);
  }
  const double r2 = dx * dx + dy * dy + dz * dz;
  double ir2 = s2 * rcp64(GUARD && j == i ? 1. : r2);
  if constexpr (GUARD) ir2 = j == i ? 0. : ir2;
  const double ir6 = ir2 * ir2 * ir2;
  const double w = ir6 * (2. * ir6 - 1.);
  const double sc = w * ir2;
  fx += sc * dx;
  fy += sc * dy;
  fz += sc * dz;
  a += ir6;
  b += w;
}

// pairs (i, j), j in [jb, je), j streamed through SGPRs (uniform scalar
// loads), four j loaded ahead of their pairs: a loop of one j waited on each
// scalar load in turn
template <bool PERIODIC, bool GUARD>
__device__ __forceinline__ void force64_range(const double4* __restrict__ x, const double4& xi,
                                              int i, int jb, int je, double L, double invL,
                                              double s2, double& fx, double& fy, double& fz,
                                              double& a, double& b)
{
  int j = __builtin_amdgcn_readfirstlane(jb);
  for (; j + 4 <= je; j += 4) {
    const double4 x0 = x[j], x1 = x[j + 1], x2 = x[j + 2], x3 = x[j + 3];
    pair64<PERIODIC, GUARD>(xi, x0, i, j, L, invL, s2, fx, fy, fz, a, b);
    pair64<PERIODIC, GUARD>(xi, x1, i, j + 1, L, invL, s2, fx, fy, fz, a, b);
    pair64<PERIODIC, GUARD>(xi, x2, i, j + 2, L, invL, s2, fx, fy, fz, a, b);
    pair64<PERIODIC, GUARD>(xi, x3, i, j + 3, L, invL, s2, fx, fy, fz, a, b);
  }
  for (; j < je; ++j) pair64<PERIODIC, GUARD>(xi, x[j], i, j, L, invL, s2, fx, fy, fz, a, b);
}

template <bool PERIODIC>
__global__ __launch_bounds__(kBlock) void k_force64(const double4* __restrict__ x, int N, int Npad,
                                                    int jchunk, double L, double invL, double s2,
                                                    double4* __restrict__ part,
                                                    double* __restrict__ apart)
{
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = (blockIdx.x * kWaves + wv) * 64;
  if (i0 >= N) return;
  const int s = blockIdx.y;
  const int j0 = s * jchunk;
  const int j1 = min(N, j0 + jchunk);
  const int i = i0 + lane;
  const double4 xi = x[min(i, N - 1)];
  double fx = 0., fy = 0., fz = 0., a = 0., b = 0.;
  // the self pair can only lie in the wave's own 64 j: the guard runs there
  const int d0 = max(j0, i0), d1 = min(j1, i0 + 64);
  if (d0 >= d1) {
    force64_range<PERIODIC, false>(x, xi, i, j0, j1, L, invL, s2, fx, fy, fz, a, b);
  } else {
    force64_range<PERIODIC, false>(x, xi, i, j0, d0, L, invL, s2, fx, fy, fz, a, b);
    force64_range<PERIODIC, true>(x, xi, i, d0, d1, L, invL, s2, fx, fy, fz, a, b);
    force64_range<PERIODIC, false>(x, xi, i, d1, j1, L, invL, s2, fx, fy, fz, a, b);
  }
  if (i < N) {
    const size_t o = (size_t)s * Npad + i;
    store_part(part + o, fx, fy, fz, b);
    store_part(apart + o, a);
  }
}

struct Sum64 {
  double fx, fy, fz, a, b;
};

__device__ __forceinline__ Sum64 sum_slices64(const Step64& a, int i)
{
  Sum64 p{0., 0., 0., 0., 0.};
  for (int s = 0; s < a.S; ++s) {
    const size_t o = (size_t)s * a.Npad + i;
    const double4 f = a.part[o];
    p.fx += f.x;
    p.fy += f.y;
    p.fz += f.z;
    p.b += f.w;
    p.a += a.apart[o];
  }
  return p;
}

// forces of particle i into f and the FP32 mirror; returns (e_i, b_i)
__device__ __forceinline__ double4 finish_force(const Step64& a, int i, double2& eb)
{
  const Sum64 p = sum_slices64(a, i);
  const double e = 0.5 * (p.b - p.a);
  const double4 f = make_double4(a.FS * p.fx, a.FS * p.fy, a.FS * p.fz, e);
  a.f[i] = f;
  a.force[i] = make_float4((float)f.x, (float)f.y, (float)f.z, (float)e);
  eb = make_double2(e, p.b);
  return f;
}

__device__ __forceinline__ void bc64(double4& x, double4& v, int bc, double L)
{
  if (bc == LJ_BC_PERIODIC) {  // MDSystem.cpp:414-424
    if (x.x < 0.) x.x += L;
    if (x.x > L) x.x -= L;
    if (x.y < 0.) x.y += L;
    if (x.y > L) x.y -= L;
    if (x.z < 0.) x.z += L;
    if (x.z > L) x.z -= L;
  } else if (bc == LJ_BC_WALL) {  // MDSystem.cpp:425-435
    if (x.x < 0. && v.x < 0.) v.x = -v.x;
    if (x.x > L && v.x > 0.) v.x = -v.x;
    if (x.y < 0. && v.y < 0.) v.y = -v.y;
    if (x.y > L && v.y > 0.) v.y = -v.y;
    if (x.z < 0. && v.z < 0.) v.z = -v.z;
    if (x.z > L && v.z > 0.) v.z = -v.z;
  }
}

__device__ __forceinline__ double sq3(const double4& v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

__device__ __forceinline__ float4 narrow(const double4& d)
{
  return make_float4((float)d.x, (float)d.y, (float)d.z, (float)d.w);
}

__global__ __launch_bounds__(kBlock) void k_reduce64(Step64 a)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[2] = {0., 0.};
  if (i < a.N) {
    double2 eb;
    (void)finish_force(a, i, eb);
    v[0] = eb.x;
    v[1] = eb.y;
  }
  const int fld[2] = {BLK_E, BLK_B};
  block_sum_store<2>(v, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

// MDSystem.cpp:447-453 (EVN) and :473-479 (TVN) in double
template <int CANON>
__global__ __launch_bounds__(kBlock) void k_drift64(Step64 a)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.N) return;
  double4 x = a.x[i];
  const double4 f = a.f[i];
  const double4 v = a.v[i];
  const double dt = a.dt;
  x.x += dt * v.x + dt * dt * f.x / 2.;
  x.y += dt * v.y + dt * dt * f.y / 2.;
  x.z += dt * v.z + dt * dt * f.z / 2.;
  a.x[i] = x;
  if constexpr (CANON == LJ_EVN) {
    a.v[i] = make_double4(v.x + dt * f.x / 2., v.y + dt * f.y / 2., v.z + dt * f.z / 2., 0.);
  } else {
    a.tf[i] = make_double4(0.5 * f.x, 0.5 * f.y, 0.5 * f.z, 0.);
  }
}

// MDSystem.cpp:458-462, then the boundary conditions (:406-436, called at :578)
__global__ __launch_bounds__(kBlock) void k_kick64(Step64 a)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v3[3] = {0., 0., 0.};
  if (i < a.N) {
    double2 eb;
    const double4 f = finish_force(a, i, eb);
    double4 v = a.v[i];
    double4 x = a.x[i];
    const double dt = a.dt;
    v.x += dt * f.x / 2.;
    v.y += dt * f.y / 2.;
    v.z += dt * f.z / 2.;
    bc64(x, v, a.bc, a.L);
    if (a.bc == LJ_BC_PERIODIC) a.x[i] = x;
    a.v[i] = v;
    a.pos[i] = narrow(x);
    a.vel[i] = narrow(v);
    v3[0] = sq3(v);
    v3[1] = eb.x;
    v3[2] = eb.y;
  }
  const int fld[3] = {BLK_K2, BLK_E, BLK_B};
  block_sum_store<3>(v3, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

// MDSystem.cpp:484-495: the averaged force and the half-step velocities
__global__ __launch_bounds__(kBlock) void k_tvn64_a(Step64 a)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v3[3] = {0., 0., 0.};
  if (i < a.N) {
    double2 eb;
    const double4 f = finish_force(a, i, eb);
    double4 tf = a.tf[i];
    tf.x += 0.5 * f.x;
    tf.y += 0.5 * f.y;
    tf.z += 0.5 * f.z;
    a.tf[i] = tf;
    const double4 v = a.v[i];
    const double dt = a.dt;
    const double4 tv = make_double4(v.x + dt * tf.x / 2., v.y + dt * tf.y / 2., v.z + dt * tf.z / 2., 0.);
    v3[0] = sq3(tv);
    v3[1] = eb.x;
    v3[2] = eb.y;
  }
  const int fld[3] = {BLK_TV2, BLK_E, BLK_B};
  block_sum_store<3>(v3, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

// MDSystem.cpp:497-505 (T_kin over 3N, :370), then the boundary conditions
__global__ __launch_bounds__(kBlock) void k_tvn64_b(Step64 a, int nblk)
{
  const double tv2 = blk_total(a.blk, nblk, BLK_TV2);
  const double chi = sqrt(a.T0 / (tv2 * (1. / 3. / a.N)));
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v1[1] = {0.};
  if (i < a.N) {
    double4 v = a.v[i];
    double4 x = a.x[i];
    const double4 tf = a.tf[i];
    const double dt = a.dt;
    v.x = (2. * chi - 1.) * v.x + chi * dt * tf.x;
    v.y = (2. * chi - 1.) * v.y + chi * dt * tf.y;
    v.z = (2. * chi - 1.) * v.z + chi * dt * tf.z;
    bc64(x, v, a.bc, a.L);
    if (a.bc == LJ_BC_PERIODIC) a.x[i] = x;
    a.v[i] = v;
    a.pos[i] = narrow(x);
    a.vel[i] = narrow(v);
    v1[0] = sq3(v);
  }
  __syncthreads();
  const int fld[1] = {BLK_K2};
  block_sum_store<1>(v1, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

__global__ __launch_bounds__(kBlock) void k_widen(const float4* __restrict__ s, double4* __restrict__ d, int N)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < N) {
    const float4 p = s[i];
    d[i] = make_double4(p.x, p.y, p.z, p.w);
  }
}

__global__ __launch_bounds__(kBlock) void k_narrow(const double4* __restrict__ s, float4* __restrict__ d, int N)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < N) d[i] = narrow(s[i]);
}

__global__ __launch_bounds__(kBlock) void k_kinetic64(const double4* __restrict__ v, int N, double* blk)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double s[1] = {i < N ? sq3(v[i]) : 0.};
  const int fld[1] = {BLK_K2};
  block_sum_store<1>(s, blk + (size_t)blockIdx.x * kBlkFields, fld);
}

// RenormalizeVelocities (MDSystem.cpp:375-389) on the double velocities
__global__ __launch_bounds__(kBlock) void k_scale64(double4* __restrict__ v, float4* __restrict__ vel,
                                                    int N, const double* __restrict__ blk, int nblk,
                                                    double T0, double factor, lj_obs* obs)
{
  double Tkin = 0.;
  if (!(factor > 0.)) {
    Tkin = blk_total(blk, nblk, BLK_K2) * (1. / 3. / N);
    factor = sqrt(T0 / Tkin);
  }
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < N) {
    double4 w = v[i];
    w.x *= factor;
    w.y *= factor;
    w.z *= factor;
    v[i] = w;
    vel[i] = narrow(w);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && obs) {
    if (Tkin > 0.) {
      obs->K *= T0 / Tkin;
      obs->T = T0;
    } else {
      obs->K *= factor * factor;
      obs->T *= factor * factor;
    }
    obs->U = obs->K + obs->V;
  }
}

__global__ __launch_bounds__(kBlock) void k_bc64(double4* __restrict__ x, double4* __restrict__ v,
                                                 float4* __restrict__ pos, float4* __restrict__ vel,
                                                 int N, int bc, double L)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  double4 p = x[i], w = v[i];
  bc64(p, w, bc, L);
  x[i] = p;
  v[i] = w;
  pos[i] = narrow(p);
  vel[i] = narrow(w);
}

inline int nb(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_force64(hipStream_t st, const ForcePlan& pl, const double4* x, bool periodic,
                          double L, double sigma, double4* part, double* apart)
{
  dim3 grid(pl.gx, pl.S);
  const double s2 = sigma * sigma;
  if (periodic)
    hipLaunchKernelGGL(k_force64<true>, grid, dim3(kBlock), 0, st, x, pl.N, pl.Npad, pl.jchunk, L,
                       1. / L, s2, part, apart);
  else
    hipLaunchKernelGGL(k_force64<false>, grid, dim3(kBlock), 0, st, x, pl.N, pl.Npad, pl.jchunk,
                       L, 1. / L, s2, part, apart);
  return hipGetLastError();
}

hipError_t launch_reduce64(hipStream_t st, const Step64& a)
{
  hipLaunchKernelGGL(k_reduce64, dim3(nb(a.N)), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_drift64(hipStream_t st, const Step64& a, int canonical)
{
  if (canonical == LJ_EVN)
    hipLaunchKernelGGL(k_drift64<LJ_EVN>, dim3(nb(a.N)), dim3(kBlock), 0, st, a);
  else
    hipLaunchKernelGGL(k_drift64<LJ_TVN>, dim3(nb(a.N)), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_kick64(hipStream_t st, const Step64& a)
{
  hipLaunchKernelGGL(k_kick64, dim3(nb(a.N)), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_tvn64_a(hipStream_t st, const Step64& a)
{
  hipLaunchKernelGGL(k_tvn64_a, dim3(nb(a.N)), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_tvn64_b(hipStream_t st, const Step64& a)
{
  hipLaunchKernelGGL(k_tvn64_b, dim3(nb(a.N)), dim3(kBlock), 0, st, a, nb(a.N));
  return hipGetLastError();
}

hipError_t launch_widen(hipStream_t st, const float4* src, double4* dst, int N)
{
  hipLaunchKernelGGL(k_widen, dim3(nb(N)), dim3(kBlock), 0, st, src, dst, N);
  return hipGetLastError();
}

hipError_t launch_narrow(hipStream_t st, const double4* src, float4* dst, int N)
{
  hipLaunchKernelGGL(k_narrow, dim3(nb(N)), dim3(kBlock), 0, st, src, dst, N);
  return hipGetLastError();
}

hipError_t launch_kinetic64(hipStream_t st, const double4* v, int N, double* blk)
{
  hipLaunchKernelGGL(k_kinetic64, dim3(nb(N)), dim3(kBlock), 0, st, v, N, blk);
  return hipGetLastError();
}

hipError_t launch_scale64(hipStream_t st, double4* v, float4* vel, int N, const double* blk,
                          int nblk, double T0, double factor, lj_obs* obs)
{
  hipLaunchKernelGGL(k_scale64, dim3(nb(N)), dim3(kBlock), 0, st, v, vel, N, blk, nblk, T0, factor,
                     obs);
  return hipGetLastError();
}

hipError_t launch_bc64(hipStream_t st, double4* x, double4* v, float4* pos, float4* vel, int N,
                       int bc, double L)
{
  hipLaunchKernelGGL(k_bc64, dim3(nb(N)), dim3(kBlock), 0, st, x, v, pos, vel, N, bc, L);
  return hipGetLastError();
}

}  // namespace lj
