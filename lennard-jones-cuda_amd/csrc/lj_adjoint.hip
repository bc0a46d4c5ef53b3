// lj_adjoint.hip -- adjoint (reverse-mode) kernels of the LJ time step.
//
// One pair pass at positions x with an adjoint vector z returns, per particle,
//   hz_i = (H z)_i = sum_j [phi (z_i - z_j) + 2 phi' (r.(z_i - z_j)) r]   (H = dF/dx)
//   f_i  = sum_j phi r_ij                                              (force)
//   gw_i = dW/dx_i = sum_j 2 psi' r_ij                                (virial gradient)
// and the scalars E, W, dE/dsigma, dW/dsigma, (dF/dsigma)^T z, F^T z, with
//   phi  = 24 eps (2 s^12 r^-14 - s^6 r^-8),   phi' = d phi / d(r^2),
//   psi  = 24 eps (2 s^12 r^-12 - s^6 r^-6)    (pair virial r.f),
// for the LJ pair of the reference (MDSystem.cpp:289-298, MDSystem.cu:44-53)
// with well depth eps and diameter s (sigma).  Same decomposition as the
// force pass: i per lane, j streamed through SGPRs, j slices, fixed-order
// FP64 cross-slice sums.  In units q = (s/r)^2, q3 = q^3:
//   phi r   = (24 eps/s^2) q q3 (2 q3 - 1) r
//   2 phi' (r.z) r = (24 eps/s^2) (2/s^2) q^2 q3 (4 - 14 q3) (r.z) r
//   2 psi' r = (48 eps/s^2) q q3 (3 - 12 q3) r
//   d phi / d s = (24 eps/s^3) q q3 (24 q3 - 6)
#include "lj_adjoint.h"
#include "lj_device.h"

namespace lj {

namespace {

using namespace dev;

struct AdjAcc {
  float hx, hy, hz;  // H z
  float fx, fy, fz;  // force
  float gx, gy, gz;  // grad W
  float a, c, s;     // sum q3, sum q3^2, sum q q3 (24 q3 - 6) (d.dz)
};

__device__ __forceinline__ void adj_zero(AdjAcc& t)
{
  t.hx = t.hy = t.hz = t.fx = t.fy = t.fz = t.gx = t.gy = t.gz = t.a = t.c = t.s = 0.f;
}

__device__ __forceinline__ void adj_add(AdjAcc& h, const AdjAcc& t)
{
  h.hx += t.hx;
  h.hy += t.hy;
  h.hz += t.hz;
  h.fx += t.fx;
  h.fy += t.fy;
  h.fz += t.fz;
  h.gx += t.gx;
  h.gy += t.gy;
  h.gz += t.gz;
  h.a += t.a;
  h.c += t.c;
  h.s += t.s;
}

// K2 = 2 u^2 / s^2 (u: length of one internal position unit)
template <bool PERIODIC, bool GUARD>
__device__ __forceinline__ void adj_pair(const PosT<PERIODIC>& pi, const PosT<PERIODIC>& pj,
                                         const float4& zi, const float4& zj, int i, int j,
                                         float C1, float K2, AdjAcc& t)
{
  float dx, dy, dz;
  if constexpr (PERIODIC) {  // exact half-box ties resolved as in the force pass
    dx = mi_fixed<true>(pi.x, pj.x);
    dy = mi_fixed<true>(pi.y, pj.y);
    dz = mi_fixed<true>(pi.z, pj.z);
  } else {
    dx = pi.x - pj.x;
    dy = pi.y - pj.y;
    dz = pi.z - pj.z;
  }
  const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
  float q = __builtin_amdgcn_rcpf(r2) * C1;
  if constexpr (GUARD) q = (i == j) ? 0.f : q;
  const float q3 = q * q * q;
  const float w = q3 * __builtin_fmaf(q3, 2.f, -1.f);  // q3 (2 q3 - 1)
  const float A1 = q * w;                                // force / H z first term
  const float ex = zi.x - zj.x, ey = zi.y - zj.y, ez = zi.z - zj.z;
  const float ddz = __builtin_fmaf(dz, ez, __builtin_fmaf(dy, ey, dx * ex));
  const float A2 = (K2 * q * q) * q3 * __builtin_fmaf(q3, -14.f, 4.f);
  const float T = A2 * ddz;
  t.hx = __builtin_fmaf(A1, ex, __builtin_fmaf(T, dx, t.hx));
  t.hy = __builtin_fmaf(A1, ey, __builtin_fmaf(T, dy, t.hy));
  t.hz = __builtin_fmaf(A1, ez, __builtin_fmaf(T, dz, t.hz));
  t.fx = __builtin_fmaf(A1, dx, t.fx);
  t.fy = __builtin_fmaf(A1, dy, t.fy);
  t.fz = __builtin_fmaf(A1, dz, t.fz);
  const float G = q * q3 * __builtin_fmaf(q3, -12.f, 3.f);
  t.gx = __builtin_fmaf(G, dx, t.gx);
  t.gy = __builtin_fmaf(G, dy, t.gy);
  t.gz = __builtin_fmaf(G, dz, t.gz);
  t.a += q3;
  t.c = __builtin_fmaf(q3, q3, t.c);
  t.s = __builtin_fmaf(q * q3 * __builtin_fmaf(q3, 24.f, -6.f), ddz, t.s);
}

template <bool PERIODIC, bool GUARD>
__device__ __forceinline__ void adj_range(const uint4* __restrict__ q, const float4* __restrict__ pf,
                                          const float4* __restrict__ z, int jb, int je,
                                          const PosT<PERIODIC>& pi, const float4& zi, int i,
                                          float C1, float K2, AdjAcc& acc)
{
  for (int b0 = jb; b0 < je; b0 += kJB) {
    const int b1 = min(je, b0 + kJB);
    AdjAcc t;
    adj_zero(t);
#pragma unroll 4
    for (int j = b0; j < b1; ++j) {
      const PosT<PERIODIC> pj = load_pos<PERIODIC>(q, pf, j);
      const float4 zj = z[j];
      adj_pair<PERIODIC, GUARD>(pi, pj, zi, zj, i, j, C1, K2, t);
    }
    adj_add(acc, t);
  }
}

// grid (gx, S); parts: 3 float4 per (slice, i): (hx,hy,hz,a), (fx,fy,fz,c), (gx,gy,gz,s)
template <bool PERIODIC>
__global__ __launch_bounds__(kBlock) void k_adj_pass(const uint4* __restrict__ q,
                                                     const float4* __restrict__ pf,
                                                     const float4* __restrict__ z,
                                                     float4* __restrict__ part, int N, int Npad,
                                                     int jchunk, float C1, float K2)
{
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = (blockIdx.x * kWaves + wv) * 64;
  if (i0 >= N) return;
  const int s = blockIdx.y;
  const int j0 = s * jchunk;
  const int j1 = min(N, j0 + jchunk);
  const int i = i0 + lane;
  const int ic = min(i, N - 1);
  const PosT<PERIODIC> pi = load_pos<PERIODIC>(q, pf, ic);
  const float4 zi = z[ic];
  AdjAcc acc;
  adj_zero(acc);
  const int d0 = max(j0, i0), d1 = min(j1, i0 + 64);
  if (d0 >= d1) {
    adj_range<PERIODIC, false>(q, pf, z, j0, j1, pi, zi, i, C1, K2, acc);
  } else {
    adj_range<PERIODIC, false>(q, pf, z, j0, d0, pi, zi, i, C1, K2, acc);
    adj_range<PERIODIC, true>(q, pf, z, d0, d1, pi, zi, i, C1, K2, acc);
    adj_range<PERIODIC, false>(q, pf, z, d1, j1, pi, zi, i, C1, K2, acc);
  }
  if (i < N) {
    const size_t o = 3 * ((size_t)s * Npad + i);
    store_part(part + o, acc.hx, acc.hy, acc.hz, acc.a);
    store_part(part + o + 1, acc.fx, acc.fy, acc.fz, acc.c);
    store_part(part + o + 2, acc.gx, acc.gy, acc.gz, acc.s);
  }
}

// per particle: sum slices in FP64, scale, write hz/f/gw (double4), block
// partials of {sum (c-a), sum (2c-a), sum (24c-6a), sum s, sum f.z}
__global__ __launch_bounds__(kBlock) void k_adj_reduce(const float4* __restrict__ part,
                                                       const float4* __restrict__ z, int N, int S,
                                                       int Npad, AdjScales sc,
                                                       double4* __restrict__ hz,
                                                       double4* __restrict__ f,
                                                       double4* __restrict__ gw,
                                                       double* __restrict__ blk)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[5] = {0., 0., 0., 0., 0.};
  if (i < N) {
    double h[3] = {0., 0., 0.}, ff[3] = {0., 0., 0.}, g[3] = {0., 0., 0.}, a = 0., c = 0., s = 0.;
    for (int k = 0; k < S; ++k) {
      const size_t o = 3 * ((size_t)k * Npad + i);
      const float4 p0 = part[o], p1 = part[o + 1], p2 = part[o + 2];
      h[0] += p0.x;
      h[1] += p0.y;
      h[2] += p0.z;
      a += p0.w;
      ff[0] += p1.x;
      ff[1] += p1.y;
      ff[2] += p1.z;
      c += p1.w;
      g[0] += p2.x;
      g[1] += p2.y;
      g[2] += p2.z;
      s += p2.w;
    }
    const double4 F = make_double4(sc.f * ff[0], sc.f * ff[1], sc.f * ff[2], c - a);
    hz[i] = make_double4(sc.h * h[0], sc.h * h[1], sc.h * h[2], 0.);
    f[i] = F;
    gw[i] = make_double4(sc.g * g[0], sc.g * g[1], sc.g * g[2], 0.);
    const float4 zi = z[i];
    v[0] = c - a;
    v[1] = 2. * c - a;
    v[2] = 24. * c - 6. * a;
    v[3] = s;
    v[4] = F.x * zi.x + F.y * zi.y + F.z * zi.z;
  }
  // block totals into row blockIdx.x (5 fields of an 8-double row)
  const int fld[5] = {0, 1, 2, 3, 4};
  block_sum_store<5>(v, blk + (size_t)blockIdx.x * 8, fld);
}

__global__ __launch_bounds__(kBlock) void k_adj_finalize(const double* __restrict__ blk, int nblk,
                                                         AdjScales sc, double* __restrict__ out)
{
  double v[5] = {0., 0., 0., 0., 0.};
  for (int r = threadIdx.x; r < nblk; r += kBlock)
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] += blk[(size_t)r * 8 + k];
  block_allsum<5>(v);
  if (threadIdx.x == 0) {
    out[ADJ_E] = 2. * sc.eps * v[0];
    out[ADJ_W] = 12. * sc.eps * v[1];
    out[ADJ_E_SIGMA] = 12. * sc.eps / sc.sigma * v[1];
    out[ADJ_W_SIGMA] = 12. * sc.eps / sc.sigma * v[2];
    out[ADJ_FSIGMA_Z] = 0.5 * sc.fsig * v[3];
    out[ADJ_F_Z] = v[4];
  }
}

// ---- reverse sweep vector kernels (FP64 adjoint state) ----

// lv = g_v(M) = (cP/M) (2/3) v_M / vol;  z = (dt/2) lv;  u = 0
__global__ __launch_bounds__(kBlock) void k_adj_init(const double4* __restrict__ vel, int N, double gv,
                                                     double dt, double4* __restrict__ u,
                                                     double4* __restrict__ lv,
                                                     float4* __restrict__ z)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const double4 v = vel[i];
  const double4 l = make_double4(gv * v.x, gv * v.y, gv * v.z, 0.);
  lv[i] = l;
  u[i] = make_double4(0., 0., 0., 0.);
  z[i] = make_float4((float)(0.5 * dt * l.x), (float)(0.5 * dt * l.y), (float)(0.5 * dt * l.z), 0.f);
}

// after the pass at x_m, m >= 1:
//   u <- u + g_x(m) + H z,  g_x = (1/M)(-F + cP gradW/(3 vol))
//   w  = (dt^2/2) u + (dt/2) lv
//   lv <- dt u + lv + g_v(m-1)        (g_v(0) = 0: vel_prev may be null)
//   z  = w + (dt/2) lv  (m-1 >= 1),  z = w  (m-1 == 0)
__global__ __launch_bounds__(kBlock) void k_adj_update(int N, double invM, double cw, double gv,
                                                       double dt, int last,
                                                       const double4* __restrict__ hz,
                                                       const double4* __restrict__ f,
                                                       const double4* __restrict__ gw,
                                                       const double4* __restrict__ vel_prev,
                                                       double4* __restrict__ u,
                                                       double4* __restrict__ lv,
                                                       float4* __restrict__ z)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const double4 h = hz[i], F = f[i], G = gw[i];
  double4 U = u[i], Lv = lv[i];
  U.x += invM * (-F.x + cw * G.x) + h.x;
  U.y += invM * (-F.y + cw * G.y) + h.y;
  U.z += invM * (-F.z + cw * G.z) + h.z;
  const double hd2 = 0.5 * dt * dt, hd = 0.5 * dt;
  const double wx = hd2 * U.x + hd * Lv.x, wy = hd2 * U.y + hd * Lv.y, wz = hd2 * U.z + hd * Lv.z;
  double gx = 0., gy = 0., gz = 0.;
  if (vel_prev) {
    const double4 v = vel_prev[i];
    gx = gv * v.x;
    gy = gv * v.y;
    gz = gv * v.z;
  }
  Lv.x = dt * U.x + Lv.x + gx;
  Lv.y = dt * U.y + Lv.y + gy;
  Lv.z = dt * U.z + Lv.z + gz;
  u[i] = U;
  lv[i] = Lv;
  if (last)
    z[i] = make_float4((float)wx, (float)wy, (float)wz, 0.f);
  else
    z[i] = make_float4((float)(wx + hd * Lv.x), (float)(wy + hd * Lv.y), (float)(wz + hd * Lv.z), 0.f);
}

// after the pass at x_0: dL/dx0 = u + H z
__global__ __launch_bounds__(kBlock) void k_adj_final(int N, const double4* __restrict__ hz,
                                                      const double4* __restrict__ u,
                                                      double4* __restrict__ dx0)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const double4 h = hz[i], U = u[i];
  dx0[i] = make_double4(U.x + h.x, U.y + h.y, U.z + h.z, 0.);
}

// theta accumulation on the device: g[0] = dL/deps, g[1] = dL/dsigma
//   += (dF/deps)^T z + (dF/dsigma)^T z  (+ direct terms when direct != 0)
__global__ void k_adj_theta(const double* __restrict__ sc, double eps, double invM, double cw3,
                            int direct, double* __restrict__ g)
{
  if (threadIdx.x != 0) return;
  g[0] += sc[ADJ_F_Z] / eps;
  g[1] += sc[ADJ_FSIGMA_Z];
  if (direct) {
    g[0] += invM * (sc[ADJ_E] / eps + cw3 * sc[ADJ_W] / eps);
    g[1] += invM * (sc[ADJ_E_SIGMA] + cw3 * sc[ADJ_W_SIGMA]);
  }
}

__global__ void k_adj_zero2(double* g)
{
  if (threadIdx.x < 2) g[threadIdx.x] = 0.;
}

// ---- FP64-state trajectory of the differentiable path ----
// Positions, velocities and forces are kept in double; the pair pass is the
// FP32 force kernel on fixed-point coordinates taken from the double
// positions (periodic) or on their float rounding (open box).

__global__ __launch_bounds__(kBlock) void k_dstate_init(const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel, int N,
                                                        double4* __restrict__ xd,
                                                        double4* __restrict__ vd)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const float4 p = pos[i], v = vel[i];
  xd[i] = make_double4(p.x, p.y, p.z, 0.);
  vd[i] = make_double4(v.x, v.y, v.z, 0.);
}

__device__ __forceinline__ unsigned fixqd(double x, double qs)
{
  return (unsigned)(long long)__builtin_rint(x * qs);
}

// internal coordinates of double positions: fixed point (qs > 0) or float
__global__ __launch_bounds__(kBlock) void k_dprep(const double4* __restrict__ xd, int N, double qs,
                                                  uint4* __restrict__ q, float4* __restrict__ xf)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const double4 x = xd[i];
  if (qs > 0.) q[i] = make_uint4(fixqd(x.x, qs), fixqd(x.y, qs), fixqd(x.z, qs), 0u);
  else xf[i] = make_float4((float)x.x, (float)x.y, (float)x.z, 0.f);
}

// x += dt v + dt^2/2 f;  v += dt/2 f  (double), then internal coordinates
__global__ __launch_bounds__(kBlock) void k_ddrift(double4* __restrict__ xd, double4* __restrict__ vd,
                                                   const double4* __restrict__ fd, int N, double dt,
                                                   double qs, uint4* __restrict__ q,
                                                   float4* __restrict__ xf)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  double4 x = xd[i], v = vd[i];
  const double4 f = fd[i];
  const double h = 0.5 * dt, h2 = 0.5 * dt * dt;
  x.x += dt * v.x + h2 * f.x;
  x.y += dt * v.y + h2 * f.y;
  x.z += dt * v.z + h2 * f.z;
  v.x += h * f.x;
  v.y += h * f.y;
  v.z += h * f.z;
  xd[i] = x;
  vd[i] = v;
  if (qs > 0.) q[i] = make_uint4(fixqd(x.x, qs), fixqd(x.y, qs), fixqd(x.z, qs), 0u);
  else xf[i] = make_float4((float)x.x, (float)x.y, (float)x.z, 0.f);
}

// f = FS * sum of slices (double); kick (dt != 0): v += dt/2 f; block partials
// {e_i, b_i, |v|^2} into rows of 8 doubles
__global__ __launch_bounds__(kBlock) void k_dkick(const float4* __restrict__ fpart,
                                                  const float* __restrict__ apart, int N, int S,
                                                  int Npad, double FS, double dt,
                                                  double4* __restrict__ fd, double4* __restrict__ vd,
                                                  double* __restrict__ blk)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v3[3] = {0., 0., 0.};
  if (i < N) {
    const Partial p = sum_slices(fpart, apart, i, S, Npad);
    const double4 f = make_double4(FS * p.fx, FS * p.fy, FS * p.fz, 0.5 * (p.b - p.a));
    fd[i] = f;
    double4 v = vd[i];
    const double h = 0.5 * dt;
    v.x += h * f.x;
    v.y += h * f.y;
    v.z += h * f.z;
    vd[i] = v;
    v3[0] = f.w;
    v3[1] = p.b;
    v3[2] = v.x * v.x + v.y * v.y + v.z * v.z;
  }
  const int fld[3] = {0, 1, 2};
  block_sum_store<3>(v3, blk + (size_t)blockIdx.x * 8, fld);
}

// E = 2 eps sum e, W = 12 eps sum b, P = (W/3 + sum v^2 / 3) / vol;
// acc[0] += E, acc[1] += P (accumulate != 0), acc[2] = E, acc[3] = P
__global__ __launch_bounds__(kBlock) void k_dfinal(const double* __restrict__ blk, int nblk,
                                                   double eps, double vol, int accumulate,
                                                   double* __restrict__ acc)
{
  double v[3] = {0., 0., 0.};
  for (int r = threadIdx.x; r < nblk; r += kBlock)
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] += blk[(size_t)r * 8 + k];
  block_allsum<3>(v);
  if (threadIdx.x != 0) return;
  const double E = 2. * eps * v[0], W = 12. * eps * v[1];
  const double P = (W / 3. + v[2] / 3.) / vol;
  if (accumulate) {
    acc[0] += E;
    acc[1] += P;
  }
  acc[2] = E;
  acc[3] = P;
}

__global__ void k_dzero4(double* acc)
{
  if (threadIdx.x < 4) acc[threadIdx.x] = 0.;
}

inline int nb(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_adj_pass(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                           const uint4* q, const float4* posf, const float4* z, float4* part)
{
  const double u = ps.periodic ? ps.L / 4294967296.0 : 1.0;
  const float K2 = (float)(2. * u * u / (ps.sigma * ps.sigma));
  dim3 grid(pl.gx, pl.S);
  if (ps.periodic)
    hipLaunchKernelGGL(k_adj_pass<true>, grid, dim3(kBlock), 0, st, q, posf, z, part, pl.N, pl.Npad,
                       pl.jchunk, ps.C1, K2);
  else
    hipLaunchKernelGGL(k_adj_pass<false>, grid, dim3(kBlock), 0, st, q, posf, z, part, pl.N, pl.Npad,
                       pl.jchunk, ps.C1, K2);
  return hipGetLastError();
}

AdjScales make_adj_scales(const PairScalars& ps)
{
  const double u = ps.periodic ? ps.L / 4294967296.0 : 1.0;
  const double s2 = ps.sigma * ps.sigma;
  AdjScales sc;
  sc.eps = ps.eps;
  sc.sigma = ps.sigma;
  sc.h = 24. * ps.eps / s2;
  sc.f = 24. * ps.eps * u / s2;
  sc.g = 48. * ps.eps * u / s2;
  sc.fsig = 24. * ps.eps * u / (s2 * ps.sigma);
  return sc;
}

hipError_t launch_adj_reduce(hipStream_t st, const ForcePlan& pl, const AdjScales& sc,
                             const float4* part, const float4* z, double4* hz, double4* f,
                             double4* gw, double* blk)
{
  hipLaunchKernelGGL(k_adj_reduce, dim3(nb(pl.N)), dim3(kBlock), 0, st, part, z, pl.N, pl.S, pl.Npad,
                     sc, hz, f, gw, blk);
  return hipGetLastError();
}

hipError_t launch_adj_finalize(hipStream_t st, const double* blk, int nblk, const AdjScales& sc,
                               double* out)
{
  hipLaunchKernelGGL(k_adj_finalize, dim3(1), dim3(kBlock), 0, st, blk, nblk, sc, out);
  return hipGetLastError();
}

hipError_t launch_adj_init(hipStream_t st, const double4* vel, int N, double gv, double dt,
                           double4* u, double4* lv, float4* z)
{
  hipLaunchKernelGGL(k_adj_init, dim3(nb(N)), dim3(kBlock), 0, st, vel, N, gv, dt, u, lv, z);
  return hipGetLastError();
}

hipError_t launch_adj_update(hipStream_t st, int N, double invM, double cw, double gv, double dt,
                             int last, const double4* hz, const double4* f, const double4* gw,
                             const double4* vel_prev, double4* u, double4* lv, float4* z)
{
  hipLaunchKernelGGL(k_adj_update, dim3(nb(N)), dim3(kBlock), 0, st, N, invM, cw, gv, dt, last, hz,
                     f, gw, vel_prev, u, lv, z);
  return hipGetLastError();
}

hipError_t launch_adj_final(hipStream_t st, int N, const double4* hz, const double4* u,
                            double4* dx0)
{
  hipLaunchKernelGGL(k_adj_final, dim3(nb(N)), dim3(kBlock), 0, st, N, hz, u, dx0);
  return hipGetLastError();
}

hipError_t launch_adj_theta(hipStream_t st, const double* sc, double eps, double invM, double cw3,
                            int direct, double* g)
{
  hipLaunchKernelGGL(k_adj_theta, dim3(1), dim3(64), 0, st, sc, eps, invM, cw3, direct, g);
  return hipGetLastError();
}

hipError_t launch_adj_zero2(hipStream_t st, double* g)
{
  hipLaunchKernelGGL(k_adj_zero2, dim3(1), dim3(64), 0, st, g);
  return hipGetLastError();
}

hipError_t launch_dstate_init(hipStream_t st, const float4* pos, const float4* vel, int N,
                              double4* xd, double4* vd)
{
  hipLaunchKernelGGL(k_dstate_init, dim3(nb(N)), dim3(kBlock), 0, st, pos, vel, N, xd, vd);
  return hipGetLastError();
}

hipError_t launch_dprep(hipStream_t st, const double4* xd, int N, double qs, uint4* q, float4* xf)
{
  hipLaunchKernelGGL(k_dprep, dim3(nb(N)), dim3(kBlock), 0, st, xd, N, qs, q, xf);
  return hipGetLastError();
}

hipError_t launch_ddrift(hipStream_t st, double4* xd, double4* vd, const double4* fd, int N,
                         double dt, double qs, uint4* q, float4* xf)
{
  hipLaunchKernelGGL(k_ddrift, dim3(nb(N)), dim3(kBlock), 0, st, xd, vd, fd, N, dt, qs, q, xf);
  return hipGetLastError();
}

hipError_t launch_dkick(hipStream_t st, const ForcePlan& pl, double FS, double dt,
                        const float4* fpart, const float* apart, double4* fd, double4* vd,
                        double* blk)
{
  hipLaunchKernelGGL(k_dkick, dim3(nb(pl.N)), dim3(kBlock), 0, st, fpart, apart, pl.N, pl.S,
                     pl.Npad, FS, dt, fd, vd, blk);
  return hipGetLastError();
}

hipError_t launch_dfinal(hipStream_t st, const double* blk, int nblk, double eps, double vol,
                         int accumulate, double* acc)
{
  hipLaunchKernelGGL(k_dfinal, dim3(1), dim3(kBlock), 0, st, blk, nblk, eps, vol, accumulate, acc);
  return hipGetLastError();
}

hipError_t launch_dzero4(hipStream_t st, double* acc)
{
  hipLaunchKernelGGL(k_dzero4, dim3(1), dim3(64), 0, st, acc);
  return hipGetLastError();
}

}  // namespace lj
