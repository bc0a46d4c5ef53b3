// lj_rdf_device.h -- device arithmetic of the RDF passes (lj_kernels.hip,
// lj_rdf_cells.hip): the reference's image and bin arithmetic as step
// functions of host-bisected thresholds, and the classification of a pair
// from the fixed-point minimum image under a proven error bound.
#pragma once

#include "lj_device.h"

namespace lj {
namespace rdfdev {

using namespace dev;

__device__ __forceinline__ float mi_cpu(float r, double L, const float (&t)[4])
{
#pragma clang fp contract(off)
  // k = fast_round((float)(r / L)) (MDSystem.cpp:274-276, :732-739) as a step
  // function of r with host-computed thresholds t[m-1] = min{x : k(x) >= m}
  int k = (r >= t[0]) + (r >= t[1]) + (r >= t[2]) + (r >= t[3]);
  k -= (r <= -t[0]) + (r <= -t[1]) + (r <= -t[2]) + (r <= -t[3]);
  return (float)((double)r - (double)k * L);
}

__device__ __forceinline__ float mi_gpu(float d, float Lf, const float (&t)[4])
{
#pragma clang fp contract(off)
  // MDSystem.cu:31-33: r += -L * lround(r / L) with a correctly rounded
  // division; k = lround(r / L) as a step function of r with host-computed
  // thresholds t[m-1] = min{x : k(x) >= m}
  int k = (d >= t[0]) + (d >= t[1]) + (d >= t[2]) + (d >= t[3]);
  k -= (d <= -t[0]) + (d <= -t[1]) + (d <= -t[2]) + (d <= -t[3]);
  return d - Lf * (float)k;
}

// bin of r2 in the threshold table T (T[k] = least r^2 of bin >= k, T[0] = 0
// <= r2, and r2 < T[nbins - 1] here): the largest k < nbins with T[k] <= r2.
// The estimate (r2 / dr2, or r / dr, in float with an approximate square
// root) is off by at most one bin: its relative error is below 2^-21 on bin
// indices below 2^13, so one step down and one step up on the table give the
// bin exactly.
__device__ __forceinline__ int rdf_bin(float r2, const float* T, int nbins, int est)
{
  int k = min(max(est, 0), nbins - 1);
  k -= T[k] > r2 ? 1 : 0;
  k += T[k + 1] <= r2 ? 1 : 0;  // k + 1 <= nbins - 1 < nbins + 1 entries
  return k;
}

// minimum image for |r| < the second threshold (image index -1, 0 or 1):
// k L is exact, so one fused rounding of r - k L in double equals the
// reference's product-then-difference
__device__ __forceinline__ float mi_cpu1(float r, double L, float t0)
{
  const int k = (r >= t0 ? 1 : 0) - (r <= -t0 ? 1 : 0);
  return (float)__builtin_fma(-(double)k, L, (double)r);
}

__device__ __forceinline__ float mi_gpu1(float d, float Lf, float t0)
{
#pragma clang fp contract(off)
  const float k = (d >= t0 ? 1.f : 0.f) - (d <= -t0 ? 1.f : 0.f);
  return d - Lf * k;
}

// The reference arithmetic for one pair: image, r^2, bin (MDSystem.cpp:269-285
// for REF_CPU and RBINS, MDSystem.cu:28-42 for REF_GPU).  Returns the bin, or
// -1 for a pair that is not counted.
// the reference's r^2 of a pair without image (hard walls), as rdf_exact_bin
// computes it: float differences, then fmaf (GPU rule) or a plain float sum
template <int RULE>
__device__ __forceinline__ float rdf_wall_r2(const float4& pi, const float4& pj)
{
#pragma clang fp contract(off)
  const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
  if constexpr (RULE == LJ_RDF_REF_GPU) return __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
  else return rx * rx + ry * ry + rz * rz;
}

template <int RULE, bool PERIODIC>
__device__ __forceinline__ int rdf_exact_bin(const float4& pi, const float4& pj, double L, float Lf,
                                             const float (&thr)[4], const float* T, int nbins,
                                             float Tmax, float Tlast, float est_scale)
{
#pragma clang fp contract(off)
  float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
  if constexpr (PERIODIC) {
    // |r| < thr[1] on every axis (any position within half a box of the box):
    // the image index is -1, 0 or 1, from the first threshold alone
    const float am = fmaxf(fabsf(rx), fmaxf(fabsf(ry), fabsf(rz)));
    if constexpr (RULE == LJ_RDF_REF_GPU) {
      if (am < thr[1]) {
        rx = mi_gpu1(rx, Lf, thr[0]);
        ry = mi_gpu1(ry, Lf, thr[0]);
        rz = mi_gpu1(rz, Lf, thr[0]);
      } else {
        rx = mi_gpu(rx, Lf, thr);
        ry = mi_gpu(ry, Lf, thr);
        rz = mi_gpu(rz, Lf, thr);
      }
    } else {
      if (am < thr[1]) {
        rx = mi_cpu1(rx, L, thr[0]);
        ry = mi_cpu1(ry, L, thr[0]);
        rz = mi_cpu1(rz, L, thr[0]);
      } else {
        rx = mi_cpu(rx, L, thr);
        ry = mi_cpu(ry, L, thr);
        rz = mi_cpu(rz, L, thr);
      }
    }
  }
  float r2;
  if constexpr (RULE == LJ_RDF_REF_GPU) {
    r2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));  // MDSystem.cu:35
    if (!(r2 > 1e-6f)) return -1;                                    // MDSystem.cu:37
  } else {
    r2 = rx * rx + ry * ry + rz * rz;  // float, no contraction (MDSystem.cpp:279)
    if (!(r2 < Tmax)) return -1;       // beyond the last bin: dropped
  }
  if (r2 >= Tlast) return nbins - 1;  // the GPU rule clamps every larger r2 here
  const float e = RULE == LJ_RDF_RBINS ? __builtin_amdgcn_sqrtf(r2) * est_scale : r2 * est_scale;
  return rdf_bin(r2, T, nbins, (int)e);
}

// The error bound of the fixed-point classification.  The fixed-point r2
// (units of (L / 2^32)^2, S = 2^32 / L per length) differs from the
// reference's r^2 (float positions, the rule's image and roundings, in S^2
// units) by at most
//   delta = Kr r + K2 r2 + K0,
//   Kr = 2 sqrt(3) (2 A + 3 L) u S,  K2 = 16 u,  K0 = 3 u^2 (2 A + 5 L)^2 S^2 + Kr^2,
// u = 2^-24, r = sqrt(r2), A a bound of |x_i - x_j| per axis.  Per axis the two
// differences agree within e = u (2 A + 3 L + 2 |r_a|): the reference's
// float difference rounds at u A, its image shift at u |r_a| (and the GPU
// rule's float L is off by u L); the fixed-point difference is within L 2^-32
// of the exact one and converts at u |r_a|; where the rule and the fixed
// point take different images (|r_a| within u (A + L) of L/2), the two
// magnitudes differ by at most 2 u (A + L) + 2 L 2^-32.  Squares differ by
// e (2 |r_a| + e), sum_a |r_a| <= sqrt(3) r; the reference's r^2 rounds at
// 3 u r2, ours at 3 u r2, the interval ends and the table's units at u r2
// each (K2: 16 against 4 + 3 + 3 + 2 + 1); the Kr^2 in K0 covers evaluating
// Kr r at the fixed-point r rather than the exact one.  Each constant is
// taken 1% up.  A pair whose interval [r2 - delta, r2 + delta] lies inside
// one bin (or wholly beyond the counted range) takes that bin without the
// reference arithmetic; the others take it.
//   RBINS (edges T[k] = fl((k dr)^2)): the bin is decided from x = r / dr
// alone: k = floor(x) when x >= 1 and frac(x) lies in [m, 1 - m],
// m = es Kr + es^2 K0 + (K2 + 6 u) x (es = 1 / (dr S)): the first terms are
// delta / (2 r) in bins, the last covers T's rounding, the sqrt (1 ulp) and
// the products forming x.  No table read.
//   Other rules: k from the estimate, then T[k] <= r2 - delta and
// r2 + delta < T[k + 1] from one read of T2[k] = (T[k], T[k + 1]) S^2.
template <int RULE>
struct RdfClass {
  float Kr, K2, K0, M0, M1, es, kmax, fnb, Tend, Tlow;
  int nbins;
  double Krd, K2d, K0d, Tend_d;

  // A: the coordinate span, L the box side, qs = S, est_scale: the rule's
  // bin estimate scale in real units, Tlast / Tmax: the table's last two entries
  __device__ __forceinline__ void init(double A, double L, double qs, float est_scale, int nb,
                                       float Tlast, float Tmax)
  {
    const double S2 = qs * qs, u = 0x1p-24;
    nbins = nb;
    Krd = 1.01 * 3.4641016151377546 * (2. * A + 3. * L) * u * qs;
    K2d = 1.01 * 16. * u;
    K0d = 1.01 * (3. * u * u * (2. * A + 5. * L) * (2. * A + 5. * L) * S2 + Krd * Krd);
    Kr = (float)Krd;
    K2 = (float)K2d;
    K0 = (float)K0d;
    const double esd = RULE == LJ_RDF_RBINS ? est_scale / qs : est_scale / S2;
    es = (float)esd;
    M0 = (float)(1.01 * (esd * Krd + esd * esd * K0d));
    M1 = (float)(1.01 * (K2d + 6. * u));
    kmax = (float)(nb - 1);
    fnb = (float)nb;
    // the upper end of the counted range: dropped (REF_CPU, RBINS) or the
    // clamped last bin (REF_GPU) from there on; the GPU rule counts r2 > 1e-6 only
    Tend_d = (double)(RULE == LJ_RDF_REF_GPU ? Tlast : Tmax) * S2;
    Tend = (float)Tend_d;
    Tlow = (float)(1e-6 * S2);
  }
  // a bound X (fixed-point r2) beyond which r2 - delta >= the range's end:
  // r2 - delta(r2) grows with r2, and X = end + 2 delta(sqrt(2 end)) <= 2 end
  __device__ __forceinline__ float far_bound() const
  {
    const double dl2 = Krd * __builtin_sqrt(2. * Tend_d) + K2d * 2. * Tend_d + K0d;
    return (float)((Tend_d + 2. * dl2) * (1. + 0x1p-20));
  }
  // RBINS, from r2 alone: k (the bin if decided), dec (decided and counted in
  // bin k), gone (decided: beyond the range)
  __device__ __forceinline__ void rbins(float r2, int& k, bool& dec, bool& gone) const
  {
    const float x = __builtin_amdgcn_sqrtf(r2) * es;
    const float m = __builtin_fmaf(M1, x, M0);
    const float f = __builtin_amdgcn_fractf(x);
    k = (int)x;
    dec = (x >= 1.f) & (f >= m) & (f <= 1.f - m) & (k < nbins);
    gone = x - m >= fnb;
  }
  // the other rules: the bin estimate and interval; t = T2[k] is read between
  __device__ __forceinline__ void measure(float r2, int& k, float& lo, float& hi) const
  {
    const float r = __builtin_bit_cast(float, (__builtin_bit_cast(unsigned, r2) >> 1) + 0x1fc00000u);
    const float dl = __builtin_fmaf(Kr, r, __builtin_fmaf(K2, r2, K0));
    lo = r2 - dl;
    hi = r2 + dl;
    k = (int)__builtin_amdgcn_fmed3f(r2 * es, 0.f, kmax);
  }
  // gone: the GPU rule's clamp bin or dropped; dec: counted in bin k
  __device__ __forceinline__ void decide(int k, float lo, float hi, float2 t, bool& dec,
                                         bool& gone) const
  {
    gone = lo >= Tend;
    dec = (t.x <= lo) & (hi < t.y);
    if constexpr (RULE == LJ_RDF_REF_GPU) dec = dec & (lo > Tlow) & (k < nbins - 1);
  }
};

// the fixed-point r2 of a pair (the exact minimum image, modulo 2^32)
__device__ __forceinline__ float rdf_r2q(const uint4& qi, const uint4& qj)
{
  const float dx = (float)(int)(qi.x - qj.x), dy = (float)(int)(qi.y - qj.y),
              dz = (float)(int)(qi.z - qj.z);
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

}  // namespace rdfdev
}  // namespace lj
