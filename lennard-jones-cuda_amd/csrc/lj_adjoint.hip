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
#include <algorithm>
#include <cstdlib>

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

// The lane's particle i in the lane's own mirrored frame.  Periodic boxes:
// axis a of lane i is mirrored (s_a = -1) when qi_a < 2^31, so that the pair
// separation in that frame, s_a (qi_a - qj_a), is qj_a - qi_a there and
// qi_a - qj_a elsewhere: one xor-add, d' = (qj ^ m_a) + o_a, with m_a = ~0,
// o_a = qi_a + 1 (unmirrored) or m_a = 0, o_a = -qi_a (mirrored).  An exact
// half-box separation (d' = 2^31) then converts to -2^31 in the lane's frame,
// i.e. to -L/2 when qi > qj and +L/2 when qi < qj -- the tie rule of the force
// pass -- without a compare and select per axis; every other separation is
// the exact negation of the unmirrored one (integer negation and the
// conversion's round-to-nearest-even are odd).  The adjoint differences
// z_i - z_j are mirrored alike, s_a z_i - s_a z_j, and every product and sum
// of the pair then carries the factor s_a exactly, so the accumulated vectors
// are s_a times the unmirrored ones bit for bit; adj_unmirror takes it out.
template <bool PERIODIC>
struct AdjI;
template <>
struct AdjI<true> {
  unsigned mx, my, mz, ox, oy, oz;
  float sx, sy, sz;  // +-1
  float zx, zy, zz;  // s_a z_i
};
template <>
struct AdjI<false> {
  float x, y, z;
  float zx, zy, zz;
};

template <bool PERIODIC>
__device__ __forceinline__ AdjI<PERIODIC> adj_lane(const PosT<PERIODIC>& p, const float4& zi)
{
  AdjI<PERIODIC> I;
  if constexpr (PERIODIC) {
    const bool fx = p.x < 0x80000000u, fy = p.y < 0x80000000u, fz = p.z < 0x80000000u;
    I.mx = fx ? 0u : ~0u;
    I.my = fy ? 0u : ~0u;
    I.mz = fz ? 0u : ~0u;
    I.ox = fx ? 0u - p.x : p.x + 1u;
    I.oy = fy ? 0u - p.y : p.y + 1u;
    I.oz = fz ? 0u - p.z : p.z + 1u;
    I.sx = fx ? -1.f : 1.f;
    I.sy = fy ? -1.f : 1.f;
    I.sz = fz ? -1.f : 1.f;
    I.zx = I.sx * zi.x;
    I.zy = I.sy * zi.y;
    I.zz = I.sz * zi.z;
  } else {
    I.x = p.x;
    I.y = p.y;
    I.z = p.z;
    I.zx = zi.x;
    I.zy = zi.y;
    I.zz = zi.z;
  }
  return I;
}

// the vector sums back in the common frame; the scalar sum s carries -1/2
// (see adj_pair)
template <bool PERIODIC>
__device__ __forceinline__ void adj_unmirror(const AdjI<PERIODIC>& I, AdjAcc& t)
{
  if constexpr (PERIODIC) {
    t.hx *= I.sx;
    t.hy *= I.sy;
    t.hz *= I.sz;
    t.fx *= I.sx;
    t.fy *= I.sy;
    t.fz *= I.sz;
    t.gx *= I.sx;
    t.gy *= I.sy;
    t.gz *= I.sz;
  }
  t.s *= -2.f;
}

// K2 = 2 u^2 / s^2 (u: length of one internal position unit).  The sum s is
// accumulated as sum G (d.dz), G = q q3 (3 - 12 q3): q q3 (24 q3 - 6) = -2 G,
// and the factor -2 (exact in binary floating point at every rounding) is
// applied once, by adj_unmirror.
template <bool PERIODIC, bool GUARD>
__device__ __forceinline__ void adj_pair(const AdjI<PERIODIC>& I, const PosT<PERIODIC>& pj,
                                         const float4& zj, int i, int j, float C1, float K2,
                                         AdjAcc& t)
{
  float dx, dy, dz, ex, ey, ez;
  if constexpr (PERIODIC) {
    dx = (float)(int)((pj.x ^ I.mx) + I.ox);
    dy = (float)(int)((pj.y ^ I.my) + I.oy);
    dz = (float)(int)((pj.z ^ I.mz) + I.oz);
    ex = __builtin_fmaf(-I.sx, zj.x, I.zx);
    ey = __builtin_fmaf(-I.sy, zj.y, I.zy);
    ez = __builtin_fmaf(-I.sz, zj.z, I.zz);
  } else {
    dx = I.x - pj.x;
    dy = I.y - pj.y;
    dz = I.z - pj.z;
    ex = I.zx - zj.x;
    ey = I.zy - zj.y;
    ez = I.zz - zj.z;
  }
  const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
  float q = __builtin_amdgcn_rcpf(r2) * C1;
  if constexpr (GUARD) q = (i == j) ? 0.f : q;
  const float q3 = q * q * q;
  const float w = q3 * __builtin_fmaf(q3, 2.f, -1.f);  // q3 (2 q3 - 1)
  const float A1 = q * w;                                // force / H z first term
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
  t.s = __builtin_fmaf(G, ddz, t.s);
}

// pairs (i, j), j in [jb, je), summed in blocks of kJB j; the j data stream
// through SGPRs (uniform scalar loads).  Loading each group of four j while
// the previous group is evaluated (explicit double buffering) measured slower
// on MI355X (10.96 against 9.61 us at N = 2048): the compiler's own schedule
// of the unrolled loop is kept.
template <bool PERIODIC, bool GUARD>
__device__ __forceinline__ void adj_range(const uint4* __restrict__ q, const float4* __restrict__ pf,
                                          const float4* __restrict__ z, int jb, int je,
                                          const AdjI<PERIODIC>& I, int i, float C1, float K2,
                                          AdjAcc& acc)
{
  for (int b0 = jb; b0 < je; b0 += kJB) {
    const int b1 = min(je, b0 + kJB);
    AdjAcc t;
    adj_zero(t);
#pragma unroll 4
    for (int j = b0; j < b1; ++j) {
      const PosT<PERIODIC> pj = load_pos<PERIODIC>(q, pf, j);
      const float4 zj = z[j];
      adj_pair<PERIODIC, GUARD>(I, pj, zj, i, j, C1, K2, t);
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
  const AdjI<PERIODIC> I = adj_lane<PERIODIC>(load_pos<PERIODIC>(q, pf, ic), z[ic]);
  AdjAcc acc;
  adj_zero(acc);
  const int d0 = max(j0, i0), d1 = min(j1, i0 + 64);
  if (d0 >= d1) {
    adj_range<PERIODIC, false>(q, pf, z, j0, j1, I, i, C1, K2, acc);
  } else {
    adj_range<PERIODIC, false>(q, pf, z, j0, d0, I, i, C1, K2, acc);
    adj_range<PERIODIC, true>(q, pf, z, d0, d1, I, i, C1, K2, acc);
    adj_range<PERIODIC, false>(q, pf, z, d1, j1, I, i, C1, K2, acc);
  }
  adj_unmirror<PERIODIC>(I, acc);
  if (i < N) {
    const size_t o = 3 * ((size_t)s * Npad + i);
    store_part(part + o, acc.hx, acc.hy, acc.hz, acc.a);
    store_part(part + o + 1, acc.fx, acc.fy, acc.fz, acc.c);
    store_part(part + o + 2, acc.gx, acc.gy, acc.gz, acc.s);
  }
}

// The same pass with the j range of each slice split between the JW waves of
// a workgroup, which share one 64-particle i tile: grid (ceil(N / 64), S).
// At sizes where the plain grid gives about one wave per SIMD (N = 2048: 8 x
// 32 workgroups) the pass was latency-bound (16.7 us for 4.2e6 pairs); more
// waves per SIMD hide the rcp and scalar-load latencies (JW = 4: 9.5 us).  The waves' sums are
// combined in LDS in wave order (deterministic) and written to the same
// partial layout, so the reduction kernels are shared.
template <bool PERIODIC, int JW>
__global__ __launch_bounds__(64 * JW) void k_adj_pass_split(const uint4* __restrict__ q,
                                                           const float4* __restrict__ pf,
                                                           const float4* __restrict__ z,
                                                           float4* __restrict__ part, int N,
                                                           int Npad, int jchunk, float C1,
                                                           float K2)
{
  __shared__ float red[JW][12][64];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = blockIdx.x * 64;
  if (i0 >= N) return;  // a padding tile: the whole workgroup leaves
  const int s = blockIdx.y;
  const int js = s * jchunk;
  const int je = min(N, js + jchunk);
  const int jq = (jchunk + JW - 1) / JW;
  const int j0 = min(je, js + wv * jq);
  const int j1 = min(je, j0 + jq);
  const int i = i0 + lane;
  const int ic = min(i, N - 1);
  const AdjI<PERIODIC> I = adj_lane<PERIODIC>(load_pos<PERIODIC>(q, pf, ic), z[ic]);
  AdjAcc acc;
  adj_zero(acc);
  const int d0 = max(j0, i0), d1 = min(j1, i0 + 64);
  if (d0 >= d1) {
    adj_range<PERIODIC, false>(q, pf, z, j0, j1, I, i, C1, K2, acc);
  } else {
    adj_range<PERIODIC, false>(q, pf, z, j0, d0, I, i, C1, K2, acc);
    adj_range<PERIODIC, true>(q, pf, z, d0, d1, I, i, C1, K2, acc);
    adj_range<PERIODIC, false>(q, pf, z, d1, j1, I, i, C1, K2, acc);
  }
  adj_unmirror<PERIODIC>(I, acc);
  const float v[12] = {acc.hx, acc.hy, acc.hz, acc.a, acc.fx, acc.fy,
                       acc.fz, acc.c,  acc.gx, acc.gy, acc.gz, acc.s};
#pragma unroll
  for (int k = 0; k < 12; ++k) red[wv][k][lane] = v[k];
  __syncthreads();
  // waves 0, 1, 2 each add one float4 of the 12 sums over the JW waves, in
  // wave order, and store it
  if (wv >= 3 || i >= N) return;
  float t[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    t[c] = red[0][4 * wv + c][lane];
#pragma unroll
    for (int w = 1; w < JW; ++w) t[c] += red[w][4 * wv + c][lane];
  }
  store_part(part + 3 * ((size_t)s * Npad + i) + wv, t[0], t[1], t[2], t[3]);
}

// per particle: sum slices in FP64, scale, write hz/f/gw (double4), block
// partials of {sum (c-a), sum (2c-a), sum (24c-6a), sum s, sum f.z}.
// Four lanes share a particle (blocks of 64 particles, one row of blk each):
// lane r sums the r-th quarter of the slices in slice order with its loads in
// flight together, then a fixed butterfly adds the quarters.  One thread per
// particle left the pass bound by load latency (N = 2048: 8 blocks, 32
// dependent slice loads each).
constexpr int kAdjLanes = 4;
constexpr int kAdjRow = kBlock / kAdjLanes;  // particles per block / row
// the fixed-point coordinate of a double position (q = rint(x qs) mod 2^32)
__device__ __forceinline__ unsigned fixqd(double x, double qs)
{
  return (unsigned)(long long)__builtin_rint(x * qs);
}

// what the reverse sweep's update of particle i reads besides the pass's sums
struct AdjPre {
  double4 u, lv, vp, xn;
};
__device__ __forceinline__ void adj_preload(int i, const AdjStep& a, AdjPre& p)
{
  if (a.mode != 0) p.u = a.u[i];
  if (a.mode == 1) p.lv = a.lv[i];
  if (a.mode == 1 && a.vel_prev) p.vp = a.vel_prev[i];
  if (a.next_x) p.xn = a.next_x[i];
}
__device__ __forceinline__ void adj_update_one(int i, const double4& h, const double4& F,
                                               const double4& G, const AdjStep& a,
                                               const AdjPre& p, float4* z);

__global__ __launch_bounds__(kBlock) void k_adj_reduce(const float4* __restrict__ part,
                                                       float4* __restrict__ z, int N, int S,
                                                       int Npad, AdjScales sc,
                                                       double4* __restrict__ hz,
                                                       double4* __restrict__ f,
                                                       double4* __restrict__ gw,
                                                       double* __restrict__ blk, AdjStep step)
{
  const int r = threadIdx.x % kAdjLanes;
  const int i = blockIdx.x * kAdjRow + threadIdx.x / kAdjLanes;
  const int C = (S + kAdjLanes - 1) / kAdjLanes;
  const int s0 = i < N ? min(S, r * C) : 0;
  const int s1 = i < N ? min(S, s0 + C) : 0;
  // the particle's own loads go out ahead of the slice sums (their latencies overlap)
  const bool own = i < N && r == 0;
  const float4 zi = own ? z[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  AdjPre pre{};
  if (own) adj_preload(i, step, pre);
  double t[12] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
  constexpr int B = 4;
  for (int k0 = s0; k0 < s1; k0 += B) {
    float4 p[B][3];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int sk = min(k0 + k, s1 - 1);
      const size_t o = 3 * ((size_t)sk * Npad + i);
      p[k][0] = part[o];
      p[k][1] = part[o + 1];
      p[k][2] = part[o + 2];
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
      if (k0 + k >= s1) break;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        t[4 * c + 0] += p[k][c].x;
        t[4 * c + 1] += p[k][c].y;
        t[4 * c + 2] += p[k][c].z;
        t[4 * c + 3] += p[k][c].w;
      }
    }
  }
#pragma unroll
  for (int o = 1; o < kAdjLanes; o <<= 1)
#pragma unroll
    for (int k = 0; k < 12; ++k) t[k] += __shfl_xor(t[k], o, 64);
  double v[5] = {0., 0., 0., 0., 0.};
  if (i < N && r == 0) {
    // t: (hx, hy, hz, a), (fx, fy, fz, c), (gx, gy, gz, s)
    const double a = t[3], c = t[7], s = t[11];
    const double4 F = make_double4(sc.f * t[4], sc.f * t[5], sc.f * t[6], c - a);
    const double4 H = make_double4(sc.h * t[0], sc.h * t[1], sc.h * t[2], 0.);
    const double4 G = make_double4(sc.g * t[8], sc.g * t[9], sc.g * t[10], 0.);
    hz[i] = H;
    f[i] = F;
    gw[i] = G;
    v[0] = c - a;
    v[1] = 2. * c - a;
    v[2] = 24. * c - 6. * a;
    v[3] = s;
    v[4] = F.x * zi.x + F.y * zi.y + F.z * zi.z;
    if (step.next_x) {  // the next pass's coordinates (k_dprep)
      const double4 x = pre.xn;
      if (step.qs > 0.)
        step.q[i] = make_uint4(fixqd(x.x, step.qs), fixqd(x.y, step.qs), fixqd(x.z, step.qs), 0u);
      else
        step.xf[i] = make_float4((float)x.x, (float)x.y, (float)x.z, 0.f);
    }
    if (step.mode == 1) {
      adj_update_one(i, H, F, G, step, pre, z);
    } else if (step.mode == 2) {  // after the pass at x_0: dL/dx0 = u + H z
      const double4 U = pre.u;
      step.dx0[i] = make_double4(U.x + H.x, U.y + H.y, U.z + H.z, 0.);
    }
  }
  // block totals into row blockIdx.x (5 fields of an 8-double row)
  const int fld[5] = {0, 1, 2, 3, 4};
  if (step.rowacc)
    block_sum_store<5, kWaves, true>(v, step.rowacc + (size_t)blockIdx.x * 8, fld);
  else
    block_sum_store<5>(v, blk + (size_t)blockIdx.x * 8, fld);
}

__global__ __launch_bounds__(kBlock) void k_adj_finalize(const double* __restrict__ blk, int nblk,
                                                         AdjScales sc, double* __restrict__ out,
                                                         AdjStep step)
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
    if (step.g) {
      // theta accumulation: g[0] = dL/deps, g[1] = dL/dsigma += (dF/deps)^T z
      // + (dF/dsigma)^T z (+ the direct terms of the loss when direct != 0)
      double* g = step.g;
      g[0] += out[ADJ_F_Z] / step.eps;
      g[1] += out[ADJ_FSIGMA_Z];
      if (step.direct) {
        g[0] += step.invM * (out[ADJ_E] / step.eps + step.cw * out[ADJ_W] / step.eps);
        g[1] += step.invM * (out[ADJ_E_SIGMA] + step.cw * out[ADJ_W_SIGMA]);
      }
    }
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
// that update for particle i (H z, F, grad W of the pass at x_m)
__device__ __forceinline__ void adj_update_one(int i, const double4& h, const double4& F,
                                               const double4& G, const AdjStep& a,
                                               const AdjPre& p, float4* z)
{
  double4 U = p.u, Lv = p.lv;
  U.x += a.invM * (-F.x + a.cw * G.x) + h.x;
  U.y += a.invM * (-F.y + a.cw * G.y) + h.y;
  U.z += a.invM * (-F.z + a.cw * G.z) + h.z;
  const double hd2 = 0.5 * a.dt * a.dt, hd = 0.5 * a.dt;
  const double wx = hd2 * U.x + hd * Lv.x, wy = hd2 * U.y + hd * Lv.y, wz = hd2 * U.z + hd * Lv.z;
  double gx = 0., gy = 0., gz = 0.;
  if (a.vel_prev) {
    const double4 v = p.vp;
    gx = a.gv * v.x;
    gy = a.gv * v.y;
    gz = a.gv * v.z;
  }
  Lv.x = a.dt * U.x + Lv.x + gx;
  Lv.y = a.dt * U.y + Lv.y + gy;
  Lv.z = a.dt * U.z + Lv.z + gz;
  a.u[i] = U;
  a.lv[i] = Lv;
  if (a.last)
    z[i] = make_float4((float)wx, (float)wy, (float)wz, 0.f);
  else
    z[i] = make_float4((float)(wx + hd * Lv.x), (float)(wy + hd * Lv.y), (float)(wz + hd * Lv.z), 0.f);
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
                                                  double* __restrict__ blk,
                                                  const double4* __restrict__ xd,
                                                  double4* __restrict__ slot_x,
                                                  double4* __restrict__ slot_v)
{
  // four lanes share a particle (blocks of 64 particles), as in k_adj_reduce
  const int r = threadIdx.x % kAdjLanes;
  const int i = blockIdx.x * kAdjRow + threadIdx.x / kAdjLanes;
  const Partial p = sum_slices_lanes<kAdjLanes, true>(fpart, apart, i < N ? i : -1, S, Npad, r);
  double v3[3] = {0., 0., 0.};
  if (i < N && r == 0) {
    const double4 f = make_double4(FS * p.fx, FS * p.fy, FS * p.fz, 0.5 * (p.b - p.a));
    fd[i] = f;
    double4 v = vd[i];
    const double h = 0.5 * dt;
    v.x += h * f.x;
    v.y += h * f.y;
    v.z += h * f.z;
    vd[i] = v;
    if (slot_x) {
      slot_x[i] = xd[i];
      slot_v[i] = v;
    }
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

// the j-split passes run one 64-particle i tile per workgroup; the tile count
// is padded to a multiple of the 8 XCDs, so that tile t is written by
// workgroups of XCD t mod 8 (workgroups go to the XCDs round robin) whatever
// their slice, the XCD whose L2 serves the row-t block of the reduction
// kernels that read the tile's partials (tiles past N exit at once)
inline unsigned split_tiles(int N) { return (unsigned)((((N + 63) / 64) + 7) / 8 * 8); }

}  // namespace

// the j-split pass serves plans whose plain grid holds under four waves per
// SIMD of the device (1024 SIMDs); LJ_ADJ_SPLIT=0/1 forces it off or on
bool adj_split_waves(const ForcePlan& pl)
{
  static const char* e = std::getenv("LJ_ADJ_SPLIT");
  if (e && e[0] == '0') return false;
  if (e && e[0] == '1') return true;
  return (long)pl.gx * kWaves * pl.S < 4L * 1024;
}

// waves sharing a tile in the j-split pass (LJ_ADJ_SPLIT_WAVES=4 for four)
int adj_split_jw()
{
  static const char* e = std::getenv("LJ_ADJ_SPLIT_WAVES");
  return (e && e[0] == '4') ? 4 : 8;
}

// The force pass's sliced plan where its grid fills the device.  Otherwise
// the j-split pass: its waves split each slice, so the slices are fewer --
// about 2048 waves in all, at least 32 j per wave (N = 2048: 8 slices of 256
// j, 8 waves of 32 j per 64-particle tile, where the force plan's 32 slices
// of 64 j left 8 j per wave to amortise each wave's set-up and LDS
// reduction; measured on MI355X, the reverse sweep took 2.84 ms against 3.28
// ms, and 2.96 ms at 16 slices).  LJ_ADJ_SLICES overrides the slice count.
ForcePlan adj_force_plan(int N)
{
  const ForcePlan pl = make_force_plan(N, 0, 1);
  if (!adj_split_waves(pl)) return pl;
  const int jw = adj_split_jw();
  const int tiles = (N + 63) / 64;
  int S = std::max(1, std::min(2048 / (tiles * jw), N / (32 * jw)));
  static const char* e = std::getenv("LJ_ADJ_SLICES");
  if (e && std::atoi(e) > 0) S = std::atoi(e);
  return make_force_plan(N, S, 1);
}

hipError_t launch_adj_pass(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                           const uint4* q, const float4* posf, const float4* z, float4* part)
{
  const double u = ps.periodic ? ps.L / 4294967296.0 : 1.0;
  const float K2 = (float)(2. * u * u / (ps.sigma * ps.sigma));
  if (adj_split_waves(pl)) {
    dim3 grid(split_tiles(pl.N), pl.S);
    const bool w8 = adj_split_jw() == 8;
    if (ps.periodic) {
      if (w8)
        hipLaunchKernelGGL((k_adj_pass_split<true, 8>), grid, dim3(512), 0, st, q, posf, z, part,
                           pl.N, pl.Npad, pl.jchunk, ps.C1, K2);
      else
        hipLaunchKernelGGL((k_adj_pass_split<true, 4>), grid, dim3(256), 0, st, q, posf, z, part,
                           pl.N, pl.Npad, pl.jchunk, ps.C1, K2);
    } else {
      if (w8)
        hipLaunchKernelGGL((k_adj_pass_split<false, 8>), grid, dim3(512), 0, st, q, posf, z, part,
                           pl.N, pl.Npad, pl.jchunk, ps.C1, K2);
      else
        hipLaunchKernelGGL((k_adj_pass_split<false, 4>), grid, dim3(256), 0, st, q, posf, z, part,
                           pl.N, pl.Npad, pl.jchunk, ps.C1, K2);
    }
    return hipGetLastError();
  }
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
                             const float4* part, float4* z, double4* hz, double4* f,
                             double4* gw, double* blk, const AdjStep& step)
{
  hipLaunchKernelGGL(k_adj_reduce, dim3(adj_rows(pl.N)), dim3(kBlock), 0, st, part, z, pl.N, pl.S, pl.Npad,
                     sc, hz, f, gw, blk, step);
  return hipGetLastError();
}

hipError_t launch_adj_finalize(hipStream_t st, const double* blk, int nblk, const AdjScales& sc,
                               double* out, const AdjStep& step)
{
  hipLaunchKernelGGL(k_adj_finalize, dim3(1), dim3(kBlock), 0, st, blk, nblk, sc, out, step);
  return hipGetLastError();
}

hipError_t launch_adj_init(hipStream_t st, const double4* vel, int N, double gv, double dt,
                           double4* u, double4* lv, float4* z)
{
  hipLaunchKernelGGL(k_adj_init, dim3(nb(N)), dim3(kBlock), 0, st, vel, N, gv, dt, u, lv, z);
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
                        double* blk, const double4* xd, double4* slot_x, double4* slot_v)
{
  hipLaunchKernelGGL(k_dkick, dim3(adj_rows(pl.N)), dim3(kBlock), 0, st, fpart, apart, pl.N, pl.S,
                     pl.Npad, FS, dt, fd, vd, blk, xd, slot_x, slot_v);
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
