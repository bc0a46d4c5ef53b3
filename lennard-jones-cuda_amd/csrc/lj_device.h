// lj_device.h -- device building blocks shared by the single-box and the
// batched kernels: the LJ pair body, the SGPR-streamed j loop with optional
// TwoSum compensation, deterministic FP64 block reductions, and the
// per-particle helpers of the integrator.
#pragma once

#include "lj_internal.h"

namespace lj {
namespace dev {

constexpr int kWaves = kBlock / 64;  // waves per workgroup
constexpr int kJB = 32;              // j block of the compensated accumulation

struct Acc {
  float fx, fy, fz, a, b;
};

__device__ __forceinline__ void acc_zero(Acc& c) { c.fx = c.fy = c.fz = c.a = c.b = 0.f; }

// Knuth TwoSum: hi + lo += x without loss
__device__ __forceinline__ void two_sum(float& hi, float& lo, float x)
{
  float s = hi + x;
  float bp = s - hi;
  float err = (hi - (s - bp)) + (x - bp);
  hi = s;
  lo += err;
}

template <bool PERIODIC>
struct PosT;
template <>
struct PosT<true> {
  unsigned x, y, z;
};
template <>
struct PosT<false> {
  float x, y, z;
};

template <bool PERIODIC>
__device__ __forceinline__ PosT<PERIODIC> load_pos(const uint4* q, const float4* pf, int j)
{
  PosT<PERIODIC> p;
  if constexpr (PERIODIC) {
    uint4 v = q[j];
    p.x = v.x;
    p.y = v.y;
    p.z = v.z;
  } else {
    float4 v = pf[j];
    p.x = v.x;
    p.y = v.y;
    p.z = v.z;
  }
  return p;
}

// Minimum-image difference of two fixed-point coordinates: the modular
// difference.  A separation of exactly half the box (0x80000000) is its own
// negative; TIES resolves it as the reference does (fast_round / lround round
// half away from zero, MDSystem.cpp:732-739, MDSystem.cu:31-33): +L/2 maps to
// -L/2 when qi > qj and -L/2 to +L/2 when qi < qj, so the two orderings of the
// pair take opposite images.
template <bool TIES>
__device__ __forceinline__ float mi_fixed(unsigned qi, unsigned qj)
{
  const unsigned d = qi - qj;
  float f = (float)(int)d;
  if constexpr (TIES) f = (d == 0x80000000u && qi < qj) ? 2147483648.f : f;
  return f;
}

// GUARD: 0 no exclusion, 1 self by index (i == j), 2 pairs with r2 <= thr
template <bool PERIODIC, int GUARD, bool TIES = false>
__device__ __forceinline__ void pair(const PosT<PERIODIC>& pi, const PosT<PERIODIC>& pj, int i,
                                     int j, float C1, float thr, Acc& c)
{
  float dx, dy, dz;
  if constexpr (PERIODIC) {
    dx = mi_fixed<TIES>(pi.x, pj.x);
    dy = mi_fixed<TIES>(pi.y, pj.y);
    dz = mi_fixed<TIES>(pi.z, pj.z);
  } else {
    dx = pi.x - pj.x;
    dy = pi.y - pj.y;
    dz = pi.z - pj.z;
  }
  float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
  float ir2 = __builtin_amdgcn_rcpf(r2) * C1;  // (sigma/r)^2
  if constexpr (GUARD == 1) ir2 = (i == j) ? 0.f : ir2;
  if constexpr (GUARD == 2) ir2 = (r2 > thr) ? ir2 : 0.f;
  float ir6 = ir2 * ir2 * ir2;
  float w = ir6 * __builtin_fmaf(ir6, 2.f, -1.f);  // (s/r)^6 (2 (s/r)^6 - 1)
  float s = w * ir2;
  c.fx = __builtin_fmaf(dx, s, c.fx);
  c.fy = __builtin_fmaf(dy, s, c.fy);
  c.fz = __builtin_fmaf(dz, s, c.fz);
  c.a += ir6;
  c.b += w;
}

template <bool PERIODIC, int GUARD, int IPL, bool COMP, bool TIES = false>
__device__ __forceinline__ void j_range(const uint4* __restrict__ q, const float4* __restrict__ pf,
                                        int jb, int je, const PosT<PERIODIC> (&pi)[IPL],
                                        const int (&ii)[IPL], float C1, float thr, Acc (&hi)[IPL],
                                        Acc (&lo)[IPL])
{
  for (int b0 = jb; b0 < je; b0 += kJB) {
    const int b1 = min(je, b0 + kJB);
    Acc t[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) acc_zero(t[k]);
#pragma unroll 8
    for (int j = b0; j < b1; ++j) {
      const PosT<PERIODIC> pj = load_pos<PERIODIC>(q, pf, j);
#pragma unroll
      for (int k = 0; k < IPL; ++k) pair<PERIODIC, GUARD, TIES>(pi[k], pj, ii[k], j, C1, thr, t[k]);
    }
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
      if constexpr (COMP) {
        two_sum(hi[k].fx, lo[k].fx, t[k].fx);
        two_sum(hi[k].fy, lo[k].fy, t[k].fy);
        two_sum(hi[k].fz, lo[k].fz, t[k].fz);
        two_sum(hi[k].a, lo[k].a, t[k].a);
        two_sum(hi[k].b, lo[k].b, t[k].b);
      } else {
        hi[k].fx += t[k].fx;
        hi[k].fy += t[k].fy;
        hi[k].fz += t[k].fz;
        hi[k].a += t[k].a;
        hi[k].b += t[k].b;
      }
    }
  }
}


// ---------------------------------------------------------------------------
// deterministic block reduction of NV doubles: butterfly inside the wave, then
// the NW wave results in wave order.  Row `out` receives the block totals
// (ACCUM: they are added to the row, which only this block touches).
template <int NV, int NW = kWaves, bool ACCUM = false>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double* out, const int (&field)[NV])
{
  __shared__ double sh[NW][NV];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.;
    for (int w = 0; w < NW; ++w) s += sh[w][threadIdx.x];
    int f = 0;
#pragma unroll
    for (int k = 0; k < NV; ++k)
      if (k == (int)threadIdx.x) f = field[k];
    if constexpr (ACCUM)
      out[f] += s;
    else
      out[f] = s;
  }
}

// block-wide total of NV doubles over NW waves, result broadcast to every thread
template <int NV, int NW = kWaves>
__device__ __forceinline__ void block_allsum(double (&v)[NV])
{
  __shared__ double sh[NW][NV];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[wv][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.;
    for (int w = 0; w < NW; ++w) s += sh[w][k];
    v[k] = s;
  }
}

// sum of one field over the rows of blk in a fixed order (every block of the
// grid obtains the same bits)
__device__ __forceinline__ double blk_total(const double* blk, int nblk, int field)
{
  double v[1] = {0.};
  for (int r = threadIdx.x; r < nblk; r += kBlock) v[0] += blk[(size_t)r * kBlkFields + field];
  block_allsum<1>(v);
  return v[0];
}

struct Partial {
  double fx, fy, fz, a, b;
};

// Sum of the partial entries [s0, s1) of particle i in slot order.  The loads
// are issued 16 at a time ahead of the (sequential, fixed-order) adds: few
// threads per particle leave few waves in flight, so the sum is bound by load
// latency unless each thread keeps many loads outstanding.
// WITH_A = false: force-only partials of three floats (store_part3; apart is
// not read; a = b = 0).
template <bool WITH_A>
__device__ __forceinline__ float4 load_entry(const float4* __restrict__ fpart, size_t e)
{
  if constexpr (WITH_A) return fpart[e];
  const float* p = reinterpret_cast<const float*>(fpart) + 3 * e;
  return make_float4(p[0], p[1], p[2], 0.f);
}

template <bool WITH_A = true>
__device__ __forceinline__ Partial sum_range(const float4* __restrict__ fpart,
                                             const float* __restrict__ apart, int i, int s0,
                                             int s1, int Npad)
{
  constexpr int B = 16;
  Partial p{0., 0., 0., 0., 0.};
  int s = s0;
  for (; s + B <= s1; s += B) {
    float4 f[B];
    float a[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      f[k] = load_entry<WITH_A>(fpart, (size_t)(s + k) * Npad + i);
      a[k] = WITH_A ? apart[(size_t)(s + k) * Npad + i] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
      p.fx += f[k].x;
      p.fy += f[k].y;
      p.fz += f[k].z;
      p.b += f[k].w;
      p.a += a[k];
    }
  }
  // the last, shorter batch is loaded the same way, all loads in flight at
  // once, from slots clamped into the range (re-reading the last one) and
  // added only up to s1: one slot at a time, each load's latency would add up
  // (ranges of fewer than 16 slots: 32 slices or 40 groups over 4 lanes)
  if (s < s1) {
    float4 f[B];
    float a[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int sk = min(s + k, s1 - 1);
      f[k] = load_entry<WITH_A>(fpart, (size_t)sk * Npad + i);
      a[k] = WITH_A ? apart[(size_t)sk * Npad + i] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const bool in = s + k < s1;
      p.fx += in ? (double)f[k].x : 0.;
      p.fy += in ? (double)f[k].y : 0.;
      p.fz += in ? (double)f[k].z : 0.;
      p.b += in ? (double)f[k].w : 0.;
      p.a += in ? (double)a[k] : 0.;
    }
  }
  return p;
}

// Sum of the S partial entries of particle i in slot order.
template <bool WITH_A = true>
__device__ __forceinline__ Partial sum_slices(const float4* __restrict__ fpart,
                                              const float* __restrict__ apart, int i, int S,
                                              int Npad)
{
  return sum_range<WITH_A>(fpart, apart, i, 0, S, Npad);
}

// Sum of the S partial entries of particle i shared by R adjacent lanes
// (r = 0..R-1, R a power of two dividing 64): lane r sums the r-th of R
// contiguous slot ranges in slot order, then a butterfly over the R lanes
// adds the range sums.  a + b and b + a round alike, so every lane of the
// particle ends with the same bits.  i < 0 marks a lane without a particle
// (it contributes zeros); every lane of the wave must call this.
template <int R, bool WITH_A>
__device__ __forceinline__ Partial sum_slices_lanes(const float4* __restrict__ fpart,
                                                    const float* __restrict__ apart, int i,
                                                    int S, int Npad, int r)
{
  static_assert(R > 0 && (R & (R - 1)) == 0 && R <= 64, "R: a power of two up to 64");
  const int C = (S + R - 1) / R;
  const int s0 = i >= 0 ? min(S, r * C) : 0;
  const int s1 = i >= 0 ? min(S, s0 + C) : 0;
  Partial p = sum_range<WITH_A>(fpart, apart, i, s0, s1, Npad);
#pragma unroll
  for (int o = 1; o < R; o <<= 1) {
    p.fx += __shfl_xor(p.fx, o, 64);
    p.fy += __shfl_xor(p.fy, o, 64);
    p.fz += __shfl_xor(p.fz, o, 64);
    p.a += __shfl_xor(p.a, o, 64);
    p.b += __shfl_xor(p.b, o, 64);
  }
  return p;
}

// Partial entries are written once by a force pass and read once by the next
// kernel: non-temporal stores keep them from holding L2 lines that the writing
// kernel no longer needs (measured on the pair-once pass at N = 16384 inside
// the time step: 96.9 -> 92.4 us).
__device__ __forceinline__ void store_part(float4* p, float x, float y, float z, float w)
{
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 v = {x, y, z, w};
  __builtin_nontemporal_store(v, reinterpret_cast<f4*>(p));
}

__device__ __forceinline__ void store_part(float* p, float x) { __builtin_nontemporal_store(x, p); }

// Force-only partial entries (the pair-once time step): three floats per
// entry, entry e = s Npad + i at float 3 e of the slot buffer, a quarter fewer
// bytes written by the force pass and read by the integrator than a float4
__device__ __forceinline__ void store_part3(float4* base, size_t e, float x, float y, float z)
{
  float* p = reinterpret_cast<float*>(base) + 3 * e;
  __builtin_nontemporal_store(x, p);
  __builtin_nontemporal_store(y, p + 1);
  __builtin_nontemporal_store(z, p + 2);
}

__device__ __forceinline__ void store_part(double4* p, double x, double y, double z, double w)
{
  typedef double d4 __attribute__((ext_vector_type(4)));
  const d4 v = {x, y, z, w};
  __builtin_nontemporal_store(v, reinterpret_cast<d4*>(p));
}

__device__ __forceinline__ void store_part(double* p, double x) { __builtin_nontemporal_store(x, p); }

__device__ __forceinline__ unsigned fixq(float x, double qs)
{
  return (unsigned)(long long)__builtin_rint((double)x * qs);
}


__device__ __forceinline__ void apply_bc_one(float4& p, float4& v, int bc, double L)
{
#pragma clang fp contract(off)
  if (bc == LJ_BC_NONE) return;
  if (bc == LJ_BC_PERIODIC) {  // MDSystem.cpp:414-424
    if (p.x < 0.) p.x = (float)(p.x + L);
    if (p.x > L) p.x = (float)(p.x - L);
    if (p.y < 0.) p.y = (float)(p.y + L);
    if (p.y > L) p.y = (float)(p.y - L);
    if (p.z < 0.) p.z = (float)(p.z + L);
    if (p.z > L) p.z = (float)(p.z - L);
  } else {  // MDSystem.cpp:425-435
    if (p.x < 0. && v.x < 0) v.x = -v.x;
    if (p.x > L && v.x > 0) v.x = -v.x;
    if (p.y < 0. && v.y < 0) v.y = -v.y;
    if (p.y > L && v.y > 0) v.y = -v.y;
    if (p.z < 0. && v.z < 0) v.z = -v.z;
    if (p.z > L && v.z > 0) v.z = -v.z;
  }
}

__device__ __forceinline__ float kin2(const float4& v)
{
#pragma clang fp contract(off)
  float s = v.x * v.x + v.y * v.y + v.z * v.z;
  return s;
}

// Observables of a step from its block rows: CalculateParameters
// (MDSystem.cpp:325-359).  Run by one block of NW waves (the stand-alone
// finaliser, or block 0 of the next step's first integrator kernel); the
// reduction tree depends on NW only, so both give the same bits for NW = 4.
// force_terms: V and W from the block rows' E/B fields plus the per-workgroup
// totals of the force pass (a.ew, a.nitems); advance: t and the step counter
// move on and the history row of the step is written.
template <int NW>
__device__ __forceinline__ void step_observables(const StepArgs& a, int nrow, bool advance,
                                                 bool force_terms)
{
  constexpr int NT = 64 * NW;
  double v[3] = {0., 0., 0.};
  // unrolled so that each thread's loads are in flight together (this block
  // runs beside the particle blocks of an integrator kernel and should not
  // outlast them)
#pragma unroll 4
  for (int r = threadIdx.x; r < nrow; r += NT) {
    v[0] += a.blk[(size_t)r * kBlkFields + BLK_K2];
    v[1] += a.blk[(size_t)r * kBlkFields + BLK_E];
    v[2] += a.blk[(size_t)r * kBlkFields + BLK_B];
  }
  if (force_terms)  // per-workgroup totals of the time step's force pass
#pragma unroll 8
    for (int r = threadIdx.x; r < a.nitems; r += NT) {
      v[1] += a.ew[2 * (size_t)r];
      v[2] += a.ew[2 * (size_t)r + 1];
    }
  block_allsum<3, NW>(v);
  if (threadIdx.x != 0) return;
  lj_obs* o = a.obs;
  const int N = a.N;
  const double K = v[0] / 2.;
  const double T = 2. * K / 3. / N;
  if (force_terms) {
    o->V = 2. * a.eps * v[1];
    o->W = 12. * a.eps * v[2];
  }
  const double Pvir = o->W / 3.;
  const double P = (Pvir + N * T) / (N / a.rho);
  o->K = K;
  o->T = T;
  o->P = P;
  o->U = K + o->V;
  o->Pshear = 0.;
  o->av_iters += 1;
  o->av_U_tot += o->U;
  o->av_p_tot += P;
  o->av_T_tot += T;
  if (advance) {
    o->t += a.dt;
    const int64_t k = o->step - *a.hist_base;
    o->step += 1;
    if (a.hist && k >= 0 && k < a.hist_cap) {
      double* h = a.hist + 5 * k;
      h[0] = o->U;
      h[1] = T;
      h[2] = K;
      h[3] = o->V;
      h[4] = P;
    }
  }
}

// First half of the velocity-Verlet step for one particle, in the reference's
// double/float pattern: x += dt v + dt^2 f / 2 (MDSystem.cpp:447-449), then
// EVN: v += dt f / 2 (:451-453); TVN: tF = f / 2 (:477-479).
template <int CANON>
__device__ __forceinline__ void drift_one(float4& p, float4& v, float4& tf, const float4& f,
                                          double dt)
{
#pragma clang fp contract(off)
  p.x = (float)(p.x + (dt * v.x + dt * dt * f.x / 2.));
  p.y = (float)(p.y + (dt * v.y + dt * dt * f.y / 2.));
  p.z = (float)(p.z + (dt * v.z + dt * dt * f.z / 2.));
  if constexpr (CANON == LJ_EVN) {
    v.x = (float)(v.x + dt * f.x / 2.);
    v.y = (float)(v.y + dt * f.y / 2.);
    v.z = (float)(v.z + dt * f.z / 2.);
  } else {
    tf = make_float4(0.5f * f.x, 0.5f * f.y, 0.5f * f.z, 0.f);
  }
}

}  // namespace dev
}  // namespace lj
