// lj_force_n3.hip -- the force pass with every unordered pair evaluated once
// (Newton's third law), for the FP32 and compensated modes of single boxes
// (k_force_n3, k_force_n3c) and for batches (k_batch_force_n3).
//
// Physics as the reference's pair loop (MDSystem.cu:21-138, MDSystem.cpp:
// 260-309); only the order of evaluation differs.
//
// Decomposition.  The particles form groups of G = 128 W (W = 4: 512, W = 2:
// 256).  A work item is a pair of groups (g, h), h >= g, run by one workgroup
// of W x W waves: wave (wi, wj) takes the 128 particles G g + 128 wi + [0, 128)
// as i (two per lane) and the 128 particles G h + 128 wj + [0, 128) as j.
//   * The j particles sit in LDS, stored twice in a row, so that at step t
//     lane l reads j = (l + t) mod 64 of a 64-particle sub-tile with a plain
//     immediate offset; after 64 steps every lane has met every j.
//   * i sums (force, (s/r)^6, (s/r)^6 (2 (s/r)^6 - 1)) stay in registers.
//   * j sums (minus the force, (s/r)^6 - w) travel with their j particle:
//     after each step every lane passes its j sums to the next lower lane
//     (ds_bpermute), so no atomics are involved and every sum is taken in a
//     fixed order (bit-reproducible results).
//   * Diagonal items (g == h): waves with wi < wj evaluate their tile pair
//     once as above, waves with wi == wj evaluate both orders of their own
//     tile (i sums only, self pair excluded by index), waves with wi > wj
//     are idle.  The time step's variant (W = 2) instead takes the diagonal
//     items two to a workgroup with every pair once (n3_diag_pair), so that
//     every workgroup carries the same work.
//   * Split items (groups of 256, sizes whose items fill few of the resident
//     workgroups, n3_split): two workgroups per item, each taking one of the
//     two 64-particle j sub-tiles of every wave, with their own slots (S = 2
//     groups); the diagonal items stay single.
// Output slots.  Particle p of group a receives exactly one entry per group s:
// s > a from the i side of item (a, s), s < a from the j side of item (s, a),
// s == a from the diagonal item.  The partial buffers are therefore the
// [S = groups][Npad] layout of the sliced kernel and the reduction and
// integrator kernels are shared (the time step's entries are the three force
// components only, store_part3).  Entry encoding (fx, fy, fz, bf) + af with
// e = (bf - af)/2 and W = 12 eps sum bf, as for the sliced kernel:
//   i side of a once-evaluated pair: bf = 2 b, af = b + a;
//   j side:                          bf = 0,   af = sum ((s/r)^6 - w);
//   both orders (self tile):         bf = b,   af = a.
// The time-step variant's loop is 24 VALU instructions per unordered pair
// (12.4 per ordered pair over the whole kernel, measured) against 21.5 per
// ordered pair for the sliced kernel, which evaluates each pair twice.
#include <algorithm>
#include <cstdlib>

#include "lj_device.h"

namespace lj {

namespace {

using namespace dev;

constexpr int kTile = 128;  // particles per wave tile
constexpr int kAccPerWave = 2 * 4 * 64;  // sub-tiles x components x tile

// LDS of one workgroup, declared once by each kernel: the diagonal-pair
// workgroups of the time step use the same bytes with their own layout
template <int W>
struct N3Lds {
  static constexpr int kG = kTile * W;
  static constexpr int kAcc = W * W * kAccPerWave > W * 5 * kG ? W * W * kAccPerWave : W * 5 * kG;
  uint4 pos[W * 2 * 128];  // [wj][sub-tile][doubled 64]
  // j sums [wave][sub-tile][component][64], then i sums [wj][component][G]
  float acc[kAcc];
  double ew[W * W][2];
};

template <bool PERIODIC, bool TIES, bool GUARD>
__device__ __forceinline__ void terms(const PosT<PERIODIC>& pi, const uint4& pj, float C1, bool ok,
                                      float& dx, float& dy, float& dz, float& s, float& e6,
                                      float& w)
{
  if constexpr (PERIODIC) {
    dx = mi_fixed<TIES>(pi.x, pj.x);
    dy = mi_fixed<TIES>(pi.y, pj.y);
    dz = mi_fixed<TIES>(pi.z, pj.z);
  } else {
    dx = pi.x - __uint_as_float(pj.x);
    dy = pi.y - __uint_as_float(pj.y);
    dz = pi.z - __uint_as_float(pj.z);
  }
  const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
  float ir2 = __builtin_amdgcn_rcpf(r2) * C1;
  if constexpr (GUARD) ir2 = ok ? ir2 : 0.f;
  e6 = ir2 * ir2 * ir2;
  w = e6 * __builtin_fmaf(e6, 2.f, -1.f);
  s = w * ir2;
}

// rotate a value one lane down the wave: lane l receives lane (l + 1) mod 64.
// ds_bpermute does the gather in the LDS unit, off the VALU that bounds the
// pair loop: a DPP move (wave_rol:1, and row_ror alike) cost about 2.5 VALU
// issue slots here, and the kernel ran 92.6 us with it against 77.8 us with
// the permute (N = 16384, time-step variant, same box).  The LDS unit is nearly
// idle in the loop (one read per step), and the permute's latency overlaps the
// next step's i-side work.  Every lane of the wave must be active: the pair
// loops run with a full execution mask.
__device__ __forceinline__ float rot1(float v)
{
  const int lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  return __int_as_float(__builtin_amdgcn_ds_bpermute(((lane + 1) & 63) << 2, __float_as_int(v)));
}

// one tile pair evaluated once: i sums in a1/a2; j sums in registers that
// travel with their j particle (lane l holds the sums of j slot (l + t) mod 64
// at step t and passes them to lane l - 1 afterwards), so after 64 steps lane l
// holds those of slot l.  J receives (-fx, -fy, -fz, sum (s/r)^6 - w) per
// sub-tile.
// nsub: sub-tiles holding a valid j (1 when the second one is past N); st0:
// the first sub-tile taken (1 for the second half of a split item)
template <bool PERIODIC, bool TIES, bool PARTIAL, bool ENERGY>
__device__ __forceinline__ void pair_tile(const uint4* P0, int lane, const PosT<PERIODIC>& pi1,
                                          const PosT<PERIODIC>& pi2, bool v1, bool v2, float C1,
                                          int nsub, Acc& a1, Acc& a2, float (&J)[2][4], int st0 = 0)
{
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    if (st >= nsub) break;  // wave-uniform
    if (st < st0) continue;
    const uint4* P = P0 + st * 128;
    float jx = 0.f, jy = 0.f, jz = 0.f, jm = 0.f;
#pragma unroll 8
    for (int t = 0; t < 64; ++t) {
      const uint4 pj = P[lane + t];
      bool ok1 = true, ok2 = true;
      if constexpr (PARTIAL) {
        const bool jv = pj.w != 0u;
        ok1 = v1 && jv;
        ok2 = v2 && jv;
      }
      float dx1, dy1, dz1, s1, e1, w1, dx2, dy2, dz2, s2, e2, w2;
      terms<PERIODIC, TIES, PARTIAL>(pi1, pj, C1, ok1, dx1, dy1, dz1, s1, e1, w1);
      terms<PERIODIC, TIES, PARTIAL>(pi2, pj, C1, ok2, dx2, dy2, dz2, s2, e2, w2);
      a1.fx = __builtin_fmaf(dx1, s1, a1.fx);
      a1.fy = __builtin_fmaf(dy1, s1, a1.fy);
      a1.fz = __builtin_fmaf(dz1, s1, a1.fz);
      a1.a += e1;
      a1.b += w1;
      a2.fx = __builtin_fmaf(dx2, s2, a2.fx);
      a2.fy = __builtin_fmaf(dy2, s2, a2.fy);
      a2.fz = __builtin_fmaf(dz2, s2, a2.fz);
      a2.a += e2;
      a2.b += w2;
      jx = __builtin_fmaf(-dx2, s2, __builtin_fmaf(-dx1, s1, jx));
      jy = __builtin_fmaf(-dy2, s2, __builtin_fmaf(-dy1, s1, jy));
      jz = __builtin_fmaf(-dz2, s2, __builtin_fmaf(-dz1, s1, jz));
      jx = rot1(jx);
      jy = rot1(jy);
      jz = rot1(jz);
      if constexpr (ENERGY) {
        jm += (e1 - w1) + (e2 - w2);
        jm = rot1(jm);
      }
    }
    J[st][0] = jx;
    J[st][1] = jy;
    J[st][2] = jz;
    J[st][3] = jm;
  }
}

// both orders of the wave's own tile, i sums only, self pair excluded
template <bool PERIODIC, bool TIES>
__device__ __forceinline__ void self_tile(const uint4* P0, int lane, int i1, int i2, int jbase,
                                          const PosT<PERIODIC>& pi1, const PosT<PERIODIC>& pi2,
                                          bool v1, bool v2, float C1, int nsub, Acc& a1, Acc& a2,
                                          int st0 = 0)
{
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    if (st >= nsub) break;  // wave-uniform
    if (st < st0) continue;
    const uint4* P = P0 + st * 128;
#pragma unroll 8
    for (int t = 0; t < 64; ++t) {
      const int jx = lane + t;
      const uint4 pj = P[jx];
      const int j = jbase + st * 64 + (jx & 63);
      const bool jv = pj.w != 0u;
      const bool ok1 = v1 && jv && j != i1;
      const bool ok2 = v2 && jv && j != i2;
      float dx, dy, dz, s, e, w;
      terms<PERIODIC, TIES, true>(pi1, pj, C1, ok1, dx, dy, dz, s, e, w);
      a1.fx = __builtin_fmaf(dx, s, a1.fx);
      a1.fy = __builtin_fmaf(dy, s, a1.fy);
      a1.fz = __builtin_fmaf(dz, s, a1.fz);
      a1.a += e;
      a1.b += w;
      terms<PERIODIC, TIES, true>(pi2, pj, C1, ok2, dx, dy, dz, s, e, w);
      a2.fx = __builtin_fmaf(dx, s, a2.fx);
      a2.fy = __builtin_fmaf(dy, s, a2.fy);
      a2.fz = __builtin_fmaf(dz, s, a2.fz);
      a2.a += e;
      a2.b += w;
    }
  }
}

// rotate by k lanes: lane l receives lane (l + k) mod 64 (full execution mask)
__device__ __forceinline__ float rotk(float v, int k)
{
  const int lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  return __int_as_float(__builtin_amdgcn_ds_bpermute(((lane + k) & 63) << 2, __float_as_int(v)));
}

__device__ __forceinline__ void acc_two_sum(Acc& hi, Acc& lo, const Acc& x)
{
  two_sum(hi.fx, lo.fx, x.fx);
  two_sum(hi.fy, lo.fy, x.fy);
  two_sum(hi.fz, lo.fz, x.fz);
  two_sum(hi.a, lo.a, x.a);
  two_sum(hi.b, lo.b, x.b);
}

// Compensated mode (LJ_PREC_COMPENSATED): the sums of pair_tile in blocks of
// kJB = 32 pairs in FP32, each block added into a (hi, lo) pair by TwoSum, as
// the sliced kernel's j_range does.
//   i side: the lane's i meets one j per step, so a block is 32 steps; the
//   block sums go into (a1, l1) / (a2, l2) at t = 32 and t = 64.
//   j side: the travelling sums meet two i per step, so a block is 16 steps.
//   The (hi, lo) pairs of the j slots stay in their lanes between folds: at
//   the fold after step t = 16 k - 1 lane l's block holds slot (l + 16 k) mod
//   64 and its (hi, lo) slot (l + 16 (k - 1)), so the pairs are first rotated
//   by 16 lanes.  After the fold at t = 64 lane l holds slot l, as in
//   pair_tile.  The i (hi, lo) are added into a1 / a2 by the caller.
template <bool PERIODIC, bool TIES, bool PARTIAL, bool ENERGY>
__device__ __forceinline__ void pair_tile_c(const uint4* P0, int lane, const PosT<PERIODIC>& pi1,
                                            const PosT<PERIODIC>& pi2, bool v1, bool v2, float C1,
                                            int nsub, Acc& a1, Acc& a2, Acc& l1, Acc& l2,
                                            float (&J)[2][4])
{
  constexpr int NJ = ENERGY ? 4 : 3;
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    if (st >= nsub) break;  // wave-uniform
    const uint4* P = P0 + st * 128;
    float jh[4] = {0.f, 0.f, 0.f, 0.f}, jl[4] = {0.f, 0.f, 0.f, 0.f};
    Acc b1, b2;
    acc_zero(b1);
    acc_zero(b2);
#pragma unroll 1
    for (int qt = 0; qt < 4; ++qt) {  // quarters of 16 steps
      float jb[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
      for (int u = 0; u < 16; ++u) {
        const uint4 pj = P[lane + 16 * qt + u];
        bool ok1 = true, ok2 = true;
        if constexpr (PARTIAL) {
          const bool jv = pj.w != 0u;
          ok1 = v1 && jv;
          ok2 = v2 && jv;
        }
        float dx1, dy1, dz1, s1, e1, w1, dx2, dy2, dz2, s2, e2, w2;
        terms<PERIODIC, TIES, PARTIAL>(pi1, pj, C1, ok1, dx1, dy1, dz1, s1, e1, w1);
        terms<PERIODIC, TIES, PARTIAL>(pi2, pj, C1, ok2, dx2, dy2, dz2, s2, e2, w2);
        b1.fx = __builtin_fmaf(dx1, s1, b1.fx);
        b1.fy = __builtin_fmaf(dy1, s1, b1.fy);
        b1.fz = __builtin_fmaf(dz1, s1, b1.fz);
        b1.a += e1;
        b1.b += w1;
        b2.fx = __builtin_fmaf(dx2, s2, b2.fx);
        b2.fy = __builtin_fmaf(dy2, s2, b2.fy);
        b2.fz = __builtin_fmaf(dz2, s2, b2.fz);
        b2.a += e2;
        b2.b += w2;
        jb[0] = rot1(__builtin_fmaf(-dx2, s2, __builtin_fmaf(-dx1, s1, jb[0])));
        jb[1] = rot1(__builtin_fmaf(-dy2, s2, __builtin_fmaf(-dy1, s1, jb[1])));
        jb[2] = rot1(__builtin_fmaf(-dz2, s2, __builtin_fmaf(-dz1, s1, jb[2])));
        if constexpr (ENERGY) jb[3] = rot1(jb[3] + ((e1 - w1) + (e2 - w2)));
      }
#pragma unroll
      for (int c = 0; c < NJ; ++c) {
        // (the first fold rotates zeros)
        jh[c] = rotk(jh[c], 16);
        jl[c] = rotk(jl[c], 16);
        two_sum(jh[c], jl[c], jb[c]);
      }
      if (qt & 1) {  // the end of a 32-step block of the i sums
        acc_two_sum(a1, l1, b1);
        acc_two_sum(a2, l2, b2);
        acc_zero(b1);
        acc_zero(b2);
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) J[st][c] = c < NJ ? jh[c] + jl[c] : 0.f;
  }
}

// Compensated form of self_tile: the i sums in blocks of 32 steps, each
// added into (a, l) by TwoSum
template <bool PERIODIC, bool TIES>
__device__ __forceinline__ void self_tile_c(const uint4* P0, int lane, int i1, int i2, int jbase,
                                            const PosT<PERIODIC>& pi1, const PosT<PERIODIC>& pi2,
                                            bool v1, bool v2, float C1, int nsub, Acc& a1, Acc& a2,
                                            Acc& l1, Acc& l2)
{
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    if (st >= nsub) break;  // wave-uniform
    const uint4* P = P0 + st * 128;
#pragma unroll 1
    for (int half = 0; half < 2; ++half) {
      Acc b1, b2;
      acc_zero(b1);
      acc_zero(b2);
#pragma unroll 8
      for (int t = 32 * half; t < 32 * half + 32; ++t) {
        const int jx = lane + t;
        const uint4 pj = P[jx];
        const int j = jbase + st * 64 + (jx & 63);
        const bool jv = pj.w != 0u;
        const bool ok1 = v1 && jv && j != i1;
        const bool ok2 = v2 && jv && j != i2;
        float dx, dy, dz, s, e, w;
        terms<PERIODIC, TIES, true>(pi1, pj, C1, ok1, dx, dy, dz, s, e, w);
        b1.fx = __builtin_fmaf(dx, s, b1.fx);
        b1.fy = __builtin_fmaf(dy, s, b1.fy);
        b1.fz = __builtin_fmaf(dz, s, b1.fz);
        b1.a += e;
        b1.b += w;
        terms<PERIODIC, TIES, true>(pi2, pj, C1, ok2, dx, dy, dz, s, e, w);
        b2.fx = __builtin_fmaf(dx, s, b2.fx);
        b2.fy = __builtin_fmaf(dy, s, b2.fy);
        b2.fz = __builtin_fmaf(dz, s, b2.fz);
        b2.a += e;
        b2.b += w;
      }
      acc_two_sum(a1, l1, b1);
      acc_two_sum(a2, l2, b2);
    }
  }
}

__device__ __forceinline__ void encode(const Acc& a, int cls, float (&v)[5])
{
  v[0] = a.fx;
  v[1] = a.fy;
  v[2] = a.fz;
  if (cls == 1) {  // once-evaluated pairs
    v[3] = 2.f * a.b;
    v[4] = a.b + a.a;
  } else {
    v[3] = a.b;
    v[4] = a.a;
  }
}

// One step of two once-evaluated pair streams of a wave: the lane's i particle
// pi1 meets j slot (lane + t) mod 64 of sub-tile P1, pi2 that of P2 (LDS,
// doubled).  R1 / R2 hold the sums of those j slots (minus the force) and pass
// one lane down afterwards, as in pair_tile.  GUARD: pairs count only with a
// valid i (vi1, vi2) and j, and with HALF only on lanes 0-31 (the step t = 32
// of a triangle, where lanes l and l + 32 meet the same pair).
template <bool PERIODIC, bool GUARD, bool HALF>
__device__ __forceinline__ void dual_step(const uint4* P1, const uint4* P2, int lane, int t,
                                          const PosT<PERIODIC>& pi1, const PosT<PERIODIC>& pi2,
                                          bool vi1, bool vi2, float C1, float (&a1)[3],
                                          float (&a2)[3], float& ea, float& eb, float (&R1)[3],
                                          float (&R2)[3])
{
  const uint4 pj1 = P1[lane + t], pj2 = P2[lane + t];
  bool ok1 = true, ok2 = true;
  if constexpr (GUARD) {
    const bool h = !HALF || lane < 32;
    ok1 = vi1 && pj1.w != 0u && h;
    ok2 = vi2 && pj2.w != 0u && h;
  }
  float dx1, dy1, dz1, s1, e1, w1, dx2, dy2, dz2, s2, e2, w2;
  terms<PERIODIC, false, GUARD>(pi1, pj1, C1, ok1, dx1, dy1, dz1, s1, e1, w1);
  terms<PERIODIC, false, GUARD>(pi2, pj2, C1, ok2, dx2, dy2, dz2, s2, e2, w2);
  a1[0] = __builtin_fmaf(dx1, s1, a1[0]);
  a1[1] = __builtin_fmaf(dy1, s1, a1[1]);
  a1[2] = __builtin_fmaf(dz1, s1, a1[2]);
  a2[0] = __builtin_fmaf(dx2, s2, a2[0]);
  a2[1] = __builtin_fmaf(dy2, s2, a2[1]);
  a2[2] = __builtin_fmaf(dz2, s2, a2[2]);
  ea += e1 + e2;
  eb += w1 + w2;
  R1[0] = rot1(__builtin_fmaf(-dx1, s1, R1[0]));
  R1[1] = rot1(__builtin_fmaf(-dy1, s1, R1[1]));
  R1[2] = rot1(__builtin_fmaf(-dz1, s1, R1[2]));
  R2[0] = rot1(__builtin_fmaf(-dx2, s2, R2[0]));
  R2[1] = rot1(__builtin_fmaf(-dy2, s2, R2[1]));
  R2[2] = rot1(__builtin_fmaf(-dz2, s2, R2[2]));
}

// The pairs inside each 64-particle sub-tile of two sub-tiles at once
// (triangles): at step t = 1..32 lane l meets slot (l + t) mod 64, so every
// unordered pair of the sub-tile is met once (t = 32 on lanes 0-31 only).
// Afterwards lane l holds the j sums of slot (l + 33) mod 64.
// The loops of these two-stream steps advance two steps per iteration: with
// more the compiler reads further ahead from LDS and, at the 64 registers that
// eight waves per SIMD allow, spilled to scratch (the whole kernel then ran
// with a private segment).
template <bool PERIODIC, bool GUARD>
__device__ __forceinline__ void tri_pair(const uint4* P1, const uint4* P2, int lane,
                                         const PosT<PERIODIC>& pi1, const PosT<PERIODIC>& pi2,
                                         bool vi1, bool vi2, float C1, float (&a1)[3],
                                         float (&a2)[3], float& ea, float& eb, float (&R1)[3],
                                         float (&R2)[3])
{
#pragma nounroll
  for (int c = 0; c < 16; ++c) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = 2 * c + u + 1;
      if (u == 1 && c == 15)
        dual_step<PERIODIC, true, true>(P1, P2, lane, t, pi1, pi2, vi1, vi2, C1, a1, a2, ea, eb, R1,
                                                  R2);
      else
        dual_step<PERIODIC, GUARD, false>(P1, P2, lane, t, pi1, pi2, vi1, vi2, C1, a1, a2, ea, eb, R1,
                                                  R2);
    }
  }
}

// Every unordered pair of one group of 256 that the pair tile of its diagonal
// item leaves out (those inside sub-tiles 0-3 and between sub-tiles 0-1 and
// 2-3), once, by one wave; f[k] receives the sums of the lane's particle of
// sub-tile k (particles base + 64 k + [0, 64), at P + 128 k).  ea / eb: the
// wave's sums of (s/r)^6 and w.
//   1. triangles of sub-tiles 1 and 3; their i and (aligned) j sums become
//      the starting j sums of
//   2. sub-tile 0 x sub-tile 1 and sub-tile 2 x sub-tile 3 (64 steps), after
//      which lane l holds the totals of slot l of sub-tiles 1 and 3;
//   3. triangles of sub-tiles 0 and 2, adding to the i sums of step 2.
template <bool PERIODIC, bool GUARD>
__device__ __forceinline__ void tri_tiles(const uint4* P, const uint4* __restrict__ q,
                                          const float4* __restrict__ pf, int base, int N, float C1,
                                          int lane, float (&f)[4][3], float& ea, float& eb)
{
  const int i0 = base + lane;
  const bool v0 = i0 < N, v1 = i0 + 64 < N, v2 = i0 + 128 < N, v3 = i0 + 192 < N;
  {
    const PosT<PERIODIC> pi1 = load_pos<PERIODIC>(q, pf, v1 ? i0 + 64 : N - 1);
    const PosT<PERIODIC> pi3 = load_pos<PERIODIC>(q, pf, v3 ? i0 + 192 : N - 1);
    float h1[3] = {0.f, 0.f, 0.f}, h3[3] = {0.f, 0.f, 0.f};
    tri_pair<PERIODIC, GUARD>(P + 128, P + 384, lane, pi1, pi3, v1, v3, C1, f[1], f[3], ea, eb, h1,
                              h3);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      f[1][c] += rotk(h1[c], 31);
      f[3][c] += rotk(h3[c], 31);
    }
  }
  const PosT<PERIODIC> pi0 = load_pos<PERIODIC>(q, pf, v0 ? i0 : N - 1);
  const PosT<PERIODIC> pi2 = load_pos<PERIODIC>(q, pf, v2 ? i0 + 128 : N - 1);
#pragma unroll 2
  for (int t = 0; t < 64; ++t)
    dual_step<PERIODIC, GUARD, false>(P + 128, P + 384, lane, t, pi0, pi2, v0, v2, C1, f[0], f[2],
                                      ea, eb, f[1], f[3]);
  float h0[3] = {0.f, 0.f, 0.f}, h2[3] = {0.f, 0.f, 0.f};
  tri_pair<PERIODIC, GUARD>(P, P + 256, lane, pi0, pi2, v0, v2, C1, f[0], f[2], ea, eb, h0, h2);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    f[0][c] += rotk(h0[c], 31);
    f[2][c] += rotk(h2[c], 31);
  }
}

// Two diagonal items (A, A) and (B, B) of the time step's variant (W = 2) in
// one workgroup, every unordered pair of a group once: waves 0 / 1 pair the
// first 128 particles of A / B with the last 128 (pair_tile), waves 2 / 3 take
// the remaining pairs of A / B (tri_tiles) and pass their sums through LDS to
// waves 0 / 1, which add them to their own and write slot A / B of the
// group's particles.  All four waves carry the same work, so with the off-diagonal items a group count of 64 gives
// 2016 + 32 = 2048 equal workgroups, as many as are resident at once (eight
// 4-wave workgroups per CU): the 2080 items of the other decomposition leave
// 32 for a second round.  B < 0: A alone (waves 1 and 3 idle).
template <bool PERIODIC>
__device__ __forceinline__ void n3_diag_pair(N3Lds<2>& lds, const uint4* __restrict__ q,
                                             const float4* __restrict__ pf,
                                             float4* __restrict__ fpart, double* __restrict__ ew,
                                             int N, int Npad, float C1, int A, int B)
{
  static_assert(sizeof(N3Lds<2>::acc) >= 512 * sizeof(uint4), "group B's positions fit in acc");
  const int T = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(T >> 6);
  const int lane = T & 63;
  uint4* const posA = lds.pos;                            // [sub-tile][doubled 64]
  uint4* const posB = reinterpret_cast<uint4*>(lds.acc);  // same layout
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int grp = r ? B : A;
    if (grp < 0) continue;
    const int j = 256 * grp + T;
    const int jl = j < N ? j : 0;
    uint4 p;
    if constexpr (PERIODIC) {
      p = q[jl];
    } else {
      const float4 f = pf[jl];
      p = make_uint4(__float_as_uint(f.x), __float_as_uint(f.y), __float_as_uint(f.z), 0u);
    }
    p.w = j < N ? 1u : 0u;
    uint4* d = (r ? posB : posA) + (T >> 6) * 128;
    d[T & 63] = p;
    d[(T & 63) + 64] = p;
  }
  __syncthreads();

  const int grp = (wave & 1) ? B : A;
  const uint4* P = (wave & 1) ? posB : posA;
  const int base = 256 * grp;
  const bool partial = base + 256 > N;
  double se = 0., sw = 0.;
  // per lane: the sums of its particle of sub-tile k = 0..3 of the group
  float f[4][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  if (grp >= 0) {
    if (wave < 2) {
      const int i1 = base + lane, i2 = i1 + 64, jt0 = base + 128;
      if (jt0 < N) {  // the last 128 hold a particle
        const bool v1 = i1 < N, v2 = i2 < N;
        const PosT<PERIODIC> pi1 = load_pos<PERIODIC>(q, pf, v1 ? i1 : N - 1);
        const PosT<PERIODIC> pi2 = load_pos<PERIODIC>(q, pf, v2 ? i2 : N - 1);
        Acc a1, a2;
        acc_zero(a1);
        acc_zero(a2);
        float J[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
        if (partial)
          pair_tile<PERIODIC, false, true, false>(P + 256, lane, pi1, pi2, v1, v2, C1,
                                                  jt0 + 64 < N ? 2 : 1, a1, a2, J);
        else
          pair_tile<PERIODIC, false, false, false>(P + 256, lane, pi1, pi2, v1, v2, C1, 2, a1, a2,
                                                   J);
        f[0][0] = a1.fx, f[0][1] = a1.fy, f[0][2] = a1.fz;
        f[1][0] = a2.fx, f[1][1] = a2.fy, f[1][2] = a2.fz;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          f[2][c] = J[0][c];
          f[3][c] = J[1][c];
        }
        se = (double)a1.a + (double)a2.a;
        sw = (double)a1.b + (double)a2.b;
      }
    } else {
      float ea = 0.f, eb = 0.f;
      if (partial)
        tri_tiles<PERIODIC, true>(P, q, pf, base, N, C1, lane, f, ea, eb);
      else
        tri_tiles<PERIODIC, false>(P, q, pf, base, N, C1, lane, f, ea, eb);
      se = (double)ea;
      sw = (double)eb;
    }
  }
  // the triangle waves' sums to the pair-tile waves through LDS (the position
  // tiles are free once every wave is done): [group][sub-tile][component][64]
  float* const tri = reinterpret_cast<float*>(lds.pos);
  static_assert(sizeof(lds.pos) + sizeof(lds.acc) >= 2 * 4 * 3 * 64 * sizeof(float), "room");
  __syncthreads();
  if (wave >= 2 && grp >= 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) tri[(((wave & 1) * 4 + k) * 3 + c) * 64 + lane] = f[k][c];
  }
  __syncthreads();
  if (wave < 2 && grp >= 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float* t = tri + ((wave * 4 + k) * 3) * 64 + lane;
      if (base + 64 * k + lane < N)
        store_part3(fpart, (size_t)grp * Npad + base + 64 * k + lane, f[k][0] + t[0],
                    f[k][1] + t[64], f[k][2] + t[128]);
    }
  }
  // totals in the ordered-pair convention: every pair here was met once
  double te = sw - se, tw = 2. * sw;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    te += __shfl_xor(te, o, 64);
    tw += __shfl_xor(tw, o, 64);
  }
  if (lane == 0) {
    lds.ew[wave][0] = te;
    lds.ew[wave][1] = tw;
  }
  __syncthreads();
  if (T == 0) {
    ew[0] = lds.ew[0][0] + lds.ew[1][0] + lds.ew[2][0] + lds.ew[3][0];
    ew[1] = lds.ew[0][1] + lds.ew[1][1] + lds.ew[2][1] + lds.ew[3][1];
  }
}

// W x W waves per workgroup, groups of G = 128 W particles.  Eight waves per
// SIMD (at most 64 VGPRs; the loop body is unchanged by the limit, no spills).
// ENERGY = false (the time step): no per-particle energies; the partial
// entries carry the force only (three floats, store_part3; no apart), and each
// workgroup writes its
// totals ew[2 item] = sum over ordered pairs of ((s/r)^12 - (s/r)^6) and
// ew[2 item + 1] = sum of w = (s/r)^6 (2 (s/r)^6 - 1), the finaliser's E and W
// terms.  The j side then carries three sums instead of four.
// One group pair (g, h) of one box: q/pf/fpart/apart point at the box's
// arrays, ew (ENERGY = false) at the workgroup's two totals.
// SPLIT: the item is run by two workgroups, half = 0 / 1 taking the first /
// second 64-particle sub-tile of every wave's 128 j (half the pairs each), so
// that sizes whose items leave most SIMDs one to three waves run twice the
// waves.  Each half writes its own slots, 2 h + half for the i side
// and 2 g + half for the j side (zeros for the j it did not meet): every
// particle still receives exactly one entry per slot, S = 2 groups.
// COMP: the compensated mode's sums (pair_tile_c, self_tile_c; W = 2, no
// split items).
template <bool PERIODIC, bool TIES, int W, bool ENERGY, bool SPLIT = false, bool COMP = false>
__device__ __forceinline__ void n3_item(N3Lds<W>& lds, const uint4* __restrict__ q,
                                        const float4* __restrict__ pf,
                                        float4* __restrict__ fpart, float* __restrict__ apart,
                                        double* __restrict__ ew, int N, int Npad, float C1, int g,
                                        int h, int half = 0)
{
  constexpr int kG = kTile * W;
  uint4* const s_pos = lds.pos;
  float* const s_acc = lds.acc;
  const int T = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(T >> 6);
  const int lane = T & 63;
  const int wi = wave / W, wj = wave % W;
  const bool diag = g == h;
  const bool partial = kG * (h + 1) > N;

  if (T < kG) {
    const int j = kG * h + T;
    const int jl = j < N ? j : 0;
    uint4 p;
    if constexpr (PERIODIC) {
      p = q[jl];
    } else {
      const float4 f = pf[jl];
      p = make_uint4(__float_as_uint(f.x), __float_as_uint(f.y), __float_as_uint(f.z), 0u);
    }
    p.w = j < N ? 1u : 0u;
    uint4* d = s_pos + (T >> 6) * 128;  // (wj, sub-tile) = (T >> 7, (T >> 6) & 1)
    d[T & 63] = p;
    d[(T & 63) + 64] = p;
  }
  __syncthreads();

  // tiles entirely past N (the last group of a box, partial) are idle, and a
  // second sub-tile past N is skipped: wave-uniform tests
  const int jt0 = kG * h + kTile * wj, it0 = kG * g + kTile * wi;
  const bool live = jt0 < N && it0 < N;
  const int nsub = SPLIT ? min(jt0 + 64 < N ? 2 : 1, half + 1) : (jt0 + 64 < N ? 2 : 1);
  const int st0 = SPLIT ? half : 0;
  // pair tile / self tile / idle
  const int cls = !live ? 0 : (diag ? (wi < wj ? 1 : (wi == wj ? 2 : 0)) : 1);
  const int i1 = kG * g + kTile * wi + lane, i2 = i1 + 64;
  const bool v1 = i1 < N, v2 = i2 < N;
  const PosT<PERIODIC> pi1 = load_pos<PERIODIC>(q, pf, v1 ? i1 : N - 1);
  const PosT<PERIODIC> pi2 = load_pos<PERIODIC>(q, pf, v2 ? i2 : N - 1);
  Acc a1, a2;
  acc_zero(a1);
  acc_zero(a2);
  const uint4* P0 = s_pos + wj * 256;
  float J[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
  if constexpr (COMP) {
    static_assert(W == 2 && !SPLIT, "compensated items: groups of 256, not split");
    Acc l1, l2;
    acc_zero(l1);
    acc_zero(l2);
    if (cls == 1) {
      if (partial)
        pair_tile_c<PERIODIC, TIES, true, ENERGY>(P0, lane, pi1, pi2, v1, v2, C1, nsub, a1, a2, l1,
                                                  l2, J);
      else
        pair_tile_c<PERIODIC, TIES, false, ENERGY>(P0, lane, pi1, pi2, v1, v2, C1, 2, a1, a2, l1,
                                                   l2, J);
    } else if (cls == 2) {
      self_tile_c<PERIODIC, TIES>(P0, lane, i1, i2, kG * h + kTile * wj, pi1, pi2, v1, v2, C1, nsub,
                                  a1, a2, l1, l2);
    }
    a1.fx += l1.fx, a1.fy += l1.fy, a1.fz += l1.fz, a1.a += l1.a, a1.b += l1.b;
    a2.fx += l2.fx, a2.fy += l2.fy, a2.fz += l2.fz, a2.a += l2.a, a2.b += l2.b;
  } else if (cls == 1) {
    if (partial)
      pair_tile<PERIODIC, TIES, true, ENERGY>(P0, lane, pi1, pi2, v1, v2, C1, nsub, a1, a2, J, st0);
    else
      pair_tile<PERIODIC, TIES, false, ENERGY>(P0, lane, pi1, pi2, v1, v2, C1, SPLIT ? nsub : 2, a1,
                                               a2, J, st0);
  } else if (cls == 2) {
    self_tile<PERIODIC, TIES>(P0, lane, i1, i2, kG * h + kTile * wj, pi1, pi2, v1, v2, C1, nsub, a1,
                              a2, st0);
  }
  {
    float* A = s_acc + wave * kAccPerWave;
#pragma unroll
    for (int st = 0; st < 2; ++st)
#pragma unroll
      for (int c = 0; c < 4; ++c) A[(st * 4 + c) * 64 + lane] = J[st][c];
  }
  __syncthreads();

  // j sums of particle G h + T over the W i rows, in row order
  float jf[4] = {0.f, 0.f, 0.f, 0.f};
  if (T < kG) {
    const int col = T >> 7, st = (T >> 6) & 1, idx = T & 63;
    for (int r = 0; r < W; ++r) {
      const float* A = s_acc + (r * W + col) * kAccPerWave + st * 256;
#pragma unroll
      for (int c = 0; c < 4; ++c) jf[c] += A[c * 64 + idx];
    }
  }
  __syncthreads();

  if constexpr (!ENERGY) {
    // E and W totals of the item in the ordered-pair convention: pairs of a
    // pair tile count twice, those of a self tile (both orders) once
    double(*s_ew)[2] = lds.ew;
    const double m = cls == 1 ? 2. : (cls == 2 ? 1. : 0.);
    const double se = (double)a1.a + (double)a2.a, sw = (double)a1.b + (double)a2.b;
    double te = m * 0.5 * (sw - se), tw = m * sw;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      te += __shfl_xor(te, o, 64);
      tw += __shfl_xor(tw, o, 64);
    }
    if (lane == 0) {
      s_ew[wave][0] = te;
      s_ew[wave][1] = tw;
    }
    __syncthreads();
    if (T == 0) {
      double E = 0., Wv = 0.;
      for (int w = 0; w < W * W; ++w) {
        E += s_ew[w][0];
        Wv += s_ew[w][1];
      }
      ew[0] = E;
      ew[1] = Wv;
    }
  }

  // i sums through LDS: ib[wj][component][G]
  float* ib = s_acc;
  {
    float e1[5], e2[5];
    if (cls == 0) {
      for (int c = 0; c < 5; ++c) e1[c] = e2[c] = 0.f;
    } else {
      encode(a1, cls, e1);
      encode(a2, cls, e2);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      ib[(wj * 5 + c) * kG + kTile * wi + lane] = e1[c];
      ib[(wj * 5 + c) * kG + kTile * wi + lane + 64] = e2[c];
    }
  }
  __syncthreads();

  if (T < kG) {
    float v[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    for (int c2 = 0; c2 < W; ++c2)
#pragma unroll
      for (int c = 0; c < 5; ++c) v[c] += ib[(c2 * 5 + c) * kG + T];
    if (diag) {
      v[0] += jf[0];
      v[1] += jf[1];
      v[2] += jf[2];
      v[4] += jf[3];
    }
    const int i = kG * g + T;
    const size_t si = SPLIT ? 2 * h + half : h, sj = SPLIT ? 2 * g + half : g;
    if (i < N) {
      if constexpr (ENERGY) {
        store_part(fpart + si * Npad + i, v[0], v[1], v[2], v[3]);
        store_part(apart + si * Npad + i, v[4]);
      } else {
        store_part3(fpart, si * Npad + i, v[0], v[1], v[2]);
      }
    }
    if (!diag) {
      const int j = kG * h + T;
      if (j < N) {
        if constexpr (ENERGY) {
          store_part(fpart + sj * Npad + j, jf[0], jf[1], jf[2], 0.f);
          store_part(apart + sj * Npad + j, jf[3]);
        } else {
          store_part3(fpart, sj * Npad + j, jf[0], jf[1], jf[2]);
        }
      }
    }
  }
}

// SPLIT: two workgroups per item (blockIdx.x = 2 item + half), no paired
// diagonal items; the (E, W) totals of the step variant per workgroup
template <bool PERIODIC, bool TIES, int W, bool ENERGY, bool SPLIT>
__global__ __launch_bounds__(64 * W * W) __attribute__((amdgpu_waves_per_eu(8))) void k_force_n3(
    const uint4* __restrict__ q, const float4* __restrict__ pf, const int2* __restrict__ items,
    float4* __restrict__ fpart, float* __restrict__ apart, double* __restrict__ ew, int N, int Npad,
    float C1)
{
  static_assert(!SPLIT || W == 2, "split items: groups of 256");
  __shared__ N3Lds<W> lds;
  const int wg = blockIdx.x;
  const int k = SPLIT ? wg >> 1 : wg;  // item
  const int2 it = items[k];
  if constexpr (!ENERGY && W == 2 && !SPLIT) {
    if (it.y < 0) {  // diagonal pair (it.x, -2 - it.y), the second one absent when -1
      n3_diag_pair<PERIODIC>(lds, q, pf, fpart, ew + 2 * (size_t)k, N, Npad, C1, it.x, -2 - it.y);
      return;
    }
  }
  n3_item<PERIODIC, TIES, W, ENERGY, SPLIT>(lds, q, pf, fpart, apart,
                                            ENERGY ? nullptr : ew + 2 * (size_t)wg, N, Npad, C1,
                                            it.x, it.y, SPLIT ? wg & 1 : 0);
}

// Compensated mode: one workgroup per item of the unpaired list (the diagonal
// items through self_tile_c), groups of 256.  The (hi, lo) pairs of the
// compensated sums take about 26 more registers than the FP32 loop, which
// holds eight waves per SIMD at 64; the kernel runs at the occupancy its
// registers allow.
template <bool PERIODIC, bool TIES, bool ENERGY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_force_n3c(const uint4* __restrict__ q,
                                                   const float4* __restrict__ pf,
                                                   const int2* __restrict__ items,
                                                   float4* __restrict__ fpart,
                                                   float* __restrict__ apart,
                                                   double* __restrict__ ew, int N, int Npad,
                                                   float C1)
{
  __shared__ N3Lds<2> lds;
  const int2 it = items[blockIdx.x];
  n3_item<PERIODIC, TIES, 2, ENERGY, false, true>(lds, q, pf, fpart, apart,
                                                  ENERGY ? nullptr : ew + 2 * (size_t)blockIdx.x, N,
                                                  Npad, C1, it.x, it.y);
}

// batched ensemble: item = (box, g, h) over all boxes; each box's partials at
// its own rows (pbase, S = groups of 256, Npad); ENERGY: per-particle energies
// kept (single evaluations), else the time step's variant with the item's
// (E, W) totals at ew[2 e].
// PERSIST: a grid of resident workgroups drains the item queue (longest items
// first): each takes the next item from the queue head until the head passes
// nitems, which every workgroup reaches; the head is reset by the kernel that
// precedes every force pass (prep, drift, or the previous step's kick).
// Otherwise workgroup b runs item b.
template <bool PERIODIC, bool TIES, bool ENERGY>
__device__ __forceinline__ void batch_n3_item(N3Lds<2>& lds, const uint4* __restrict__ q,
                                              const float4* __restrict__ pf,
                                              float4* __restrict__ fpart, float* __restrict__ apart,
                                              const BoxParam* __restrict__ box,
                                              const BatchItem& w, double* __restrict__ ew)
{
  const BoxParam bp = box[w.box];
  n3_item<PERIODIC, TIES, 2, ENERGY>(lds, q + bp.off, pf + bp.off, fpart + bp.pbase, apart + bp.pbase,
                                     ENERGY ? nullptr : ew + 2 * (size_t)w.e, bp.N, bp.Npad, bp.C1,
                                     w.i0, w.s);
}

template <bool PERIODIC, bool TIES, bool ENERGY, bool PERSIST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_batch_force_n3(
    const uint4* __restrict__ q, const float4* __restrict__ pf, float4* __restrict__ fpart,
    float* __restrict__ apart, const BoxParam* __restrict__ box, const BatchItem* __restrict__ items,
    double* __restrict__ ew, unsigned* __restrict__ counter, int nitems)
{
  __shared__ N3Lds<2> lds;
  if constexpr (PERSIST) {
    __shared__ unsigned s_item;
    for (;;) {
      if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
      __syncthreads();  // also: every thread is done with the previous item's LDS
      const unsigned it = s_item;
      __syncthreads();  // s_item read by all before thread 0 writes it again
      if (it >= (unsigned)nitems) break;  // uniform in the workgroup
      batch_n3_item<PERIODIC, TIES, ENERGY>(lds, q, pf, fpart, apart, box, items[it], ew);
    }
  } else {
    batch_n3_item<PERIODIC, TIES, ENERGY>(lds, q, pf, fpart, apart, box, items[blockIdx.x], ew);
  }
}

}  // namespace

int n3_waves(int /*N*/)
{
  // W = 2 (groups of 256, 4-wave workgroups, up to 8 resident per CU) measured
  // 116 us at N = 16384 against 145 us for W = 4 (groups of 512, 16-wave
  // workgroups, at most 2 per CU); LJ_PAIR_ONCE_WAVES=4 selects the latter
  const char* e = std::getenv("LJ_PAIR_ONCE_WAVES");
  if (e && std::atoi(e) == 4) return 4;
  return 2;
}

// Relative pair rate of one CU holding w four-wave workgroups of the pair
// loop (one wave per SIMD each), measured on the time step's pass at
// rho* = 0.85: 1 : 2.0 : 2.15 : 2.6 for w = 1, 2, 3 and >= 4 (one wave per
// SIMD is bound by its own latency; from four on the VALU is nearly full).
double n3_cu_rate(int w) { return w <= 1 ? 1.0 : w == 2 ? 2.0 : w == 3 ? 2.15 : 2.6; }

// modelled duration of a pass of n workgroups, each `work` items of pairs,
// spread evenly over `cus` CUs: the busiest CU holds ceil(n / cus) of them
double n3_pass_time(long n, double work, int cus)
{
  const int w = (int)((n + cus - 1) / cus);
  return w * work / n3_cu_rate(w);
}

int n3_split(int W, int ng)
{
  // LJ_PAIR_ONCE_SPLIT=0 / 1 turns the split items off / on (groups of 256)
  if (W != 2) return 1;
  const char* e = std::getenv("LJ_PAIR_ONCE_SPLIT");
  if (e && (e[0] == '0' || e[0] == '1')) return e[0] == '1' ? 2 : 1;
  int cus = 256, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  // split when the model gives the split pass at least 10% less time (its
  // twice as many partial slots cost the integrator about that much).
  // Measured step times (TVN, N: split / unsplit us): 4096 21.3 / 32.6, 6144
  // 26.4 / 33.7, 8192 34.6 / 34.5, 10000 44.6 / 47.2, 11264 48.9 / 48.4.
  const long full = (long)ng * (ng + 1) / 2, paired = (long)ng * (ng - 1) / 2 + (ng + 1) / 2;
  return n3_pass_time(2 * full, 0.5, cus) < 0.9 * n3_pass_time(paired, 1.0, cus) ? 2 : 1;
}

ForcePlan make_n3_plan(int N, bool comp)
{
  ForcePlan pl;
  const int W = comp ? 2 : n3_waves(N);
  const int G = kTile * W;
  pl.N = N;
  pl.n3 = W;
  pl.comp = comp ? 1 : 0;
  pl.ng = (N + G - 1) / G;
  pl.split = comp ? 1 : n3_split(W, pl.ng);
  pl.S = pl.ng * pl.split;
  pl.jchunk = G;
  pl.Npad = (N + 63) & ~63;
  pl.nitems = pl.ng * (pl.ng + 1) / 2;
  // the time step pairs its diagonal items (n3_step_items), unless split or
  // compensated
  pl.nitems_step = W == 2 && pl.split == 1 && !comp ? pl.ng * (pl.ng - 1) / 2 + (pl.ng + 1) / 2
                                                    : pl.nitems;
  return pl;
}

bool n3_comp_enabled()
{
  // LJ_PAIR_ONCE_COMP=0 keeps the compensated mode on the sliced kernel
  const char* e = std::getenv("LJ_PAIR_ONCE_COMP");
  return !(e && e[0] == '0');
}

bool n3_eligible(int N)
{
  // LJ_PAIR_ONCE=0 / 1 turns the kernel off / on at any size (tuning)
  const char* e = std::getenv("LJ_PAIR_ONCE");
  if (e && (e[0] == '0' || e[0] == '1')) return e[0] == '1' && N >= 256;
  // measured against the sliced kernel (TVN step at rho* = 0.85, us per step,
  // with split items): N = 2560 19.2 vs 21.4, 3072 19.3 vs 22.6, 4096 21.2 vs
  // 25.1, 6144 26.4 vs 36.8.  Below N = 3584 the pass itself is slower (one
  // wave per SIMD: 10.9 against 9.4 us at 3072), the step faster for its
  // fused observables; the threshold keeps single evaluations of small boxes
  // on the sliced kernel
  return N >= 3072;
}

void n3_items(int ng, int2* out)
{
  // off-diagonal items first, the lighter diagonal ones last
  int k = 0;
  for (int g = 0; g < ng; ++g)
    for (int h = g + 1; h < ng; ++h) out[k++] = make_int2(g, h);
  for (int g = 0; g < ng; ++g) out[k++] = make_int2(g, g);
}

void n3_step_items(const ForcePlan& pl, int2* out)
{
  if (pl.nitems_step == pl.nitems) {
    n3_items(pl.ng, out);
    return;
  }
  // the diagonal items in pairs (g, g + half) encoded as (g, -2 - (g + half)),
  // a last odd one alone as (g, -1), first: when the workgroups outnumber the
  // resident slots the longer ones should start in the first round; then the
  // off-diagonal items
  const int ng = pl.ng, half = (ng + 1) / 2;
  int k = 0;
  for (int g = 0; g < half; ++g) out[k++] = make_int2(g, g + half < ng ? -2 - (g + half) : -1);
  for (int g = 0; g < ng; ++g)
    for (int h = g + 1; h < ng; ++h) out[k++] = make_int2(g, h);
}

namespace {

template <int W, bool ENERGY, bool SPLIT>
hipError_t launch_n3_w(hipStream_t st, const ForcePlan& pl, const PairScalars& ps, const uint4* q,
                       const float4* posf, float4* fpart, float* apart, double* ew, bool ties)
{
  // the time step's variant runs its own item list (diagonal items paired)
  const int nitems = ENERGY ? pl.nitems : pl.nitems_step;
  const int2* items = ENERGY ? pl.items : pl.items_step;
  dim3 grid(nitems * (SPLIT ? 2 : 1)), block(64 * W * W);
  const LaunchStamp ts = tl_force_stamp;
  if (ps.periodic) {
    if (ties)
      hipExtLaunchKernelGGL((k_force_n3<true, true, W, ENERGY, SPLIT>), grid, block, 0, st, ts.start,
                            ts.stop, 0, q, posf, items, fpart, apart, ew, pl.N, pl.Npad, ps.C1);
    else
      hipExtLaunchKernelGGL((k_force_n3<true, false, W, ENERGY, SPLIT>), grid, block, 0, st, ts.start,
                            ts.stop, 0, q, posf, items, fpart, apart, ew, pl.N, pl.Npad, ps.C1);
  } else {
    hipExtLaunchKernelGGL((k_force_n3<false, false, W, ENERGY, SPLIT>), grid, block, 0, st, ts.start,
                          ts.stop, 0, q, posf, items, fpart, apart, ew, pl.N, pl.Npad, ps.C1);
  }
  return hipGetLastError();
}

template <bool ENERGY>
hipError_t launch_n3_c(hipStream_t st, const ForcePlan& pl, const PairScalars& ps, const uint4* q,
                       const float4* posf, float4* fpart, float* apart, double* ew, bool ties)
{
  // both variants run the unpaired item list (items_step equals it)
  const int nitems = ENERGY ? pl.nitems : pl.nitems_step;
  const int2* items = ENERGY ? pl.items : pl.items_step;
  dim3 grid(nitems), block(256);
  const LaunchStamp ts = tl_force_stamp;
  if (ps.periodic) {
    if (ties)
      hipExtLaunchKernelGGL((k_force_n3c<true, true, ENERGY>), grid, block, 0, st, ts.start, ts.stop,
                            0, q, posf, items, fpart, apart, ew, pl.N, pl.Npad, ps.C1);
    else
      hipExtLaunchKernelGGL((k_force_n3c<true, false, ENERGY>), grid, block, 0, st, ts.start, ts.stop,
                            0, q, posf, items, fpart, apart, ew, pl.N, pl.Npad, ps.C1);
  } else {
    hipExtLaunchKernelGGL((k_force_n3c<false, false, ENERGY>), grid, block, 0, st, ts.start, ts.stop,
                          0, q, posf, items, fpart, apart, ew, pl.N, pl.Npad, ps.C1);
  }
  return hipGetLastError();
}

}  // namespace

template <bool PERIODIC, bool TIES, bool ENERGY>
void launch_batch_n3_t(hipStream_t st, const BatchArgs& a)
{
  const int pg = a.persistent_grid > 0 ? std::min(a.persistent_grid, a.nitems) : 0;
  if (pg > 0)
    hipLaunchKernelGGL((k_batch_force_n3<PERIODIC, TIES, ENERGY, true>), dim3(pg), dim3(256), 0, st,
                       a.q, a.pos, a.fpart, a.apart, a.box, a.items, a.ew, a.counter, a.nitems);
  else
    hipLaunchKernelGGL((k_batch_force_n3<PERIODIC, TIES, ENERGY, false>), dim3(a.nitems),
                       dim3(256), 0, st, a.q, a.pos, a.fpart, a.apart, a.box, a.items, a.ew,
                       a.counter, a.nitems);
}

hipError_t launch_batch_force_n3(hipStream_t st, const BatchArgs& a, bool periodic, bool ties)
{
  if (a.lean) {  // the time step (no exact-tie rule there)
    if (ties || !a.ew) return hipErrorInvalidValue;
    if (periodic) launch_batch_n3_t<true, false, false>(st, a);
    else launch_batch_n3_t<false, false, false>(st, a);
    return hipGetLastError();
  }
  if (periodic) {
    if (ties) launch_batch_n3_t<true, true, true>(st, a);
    else launch_batch_n3_t<true, false, true>(st, a);
  } else {
    launch_batch_n3_t<false, false, true>(st, a);
  }
  return hipGetLastError();
}

hipError_t launch_force_n3_step_items(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                                      const uint4* q, const float4* posf, float4* fpart, double* ew,
                                      int first, int count)
{
  // the step variant's items [first, first + count), each writing its own
  // partial entries and (E, W) totals: one GPU's share of a pass whose items
  // are dealt out (W = 2, diagonal items paired, not split)
  if (pl.n3 != 2 || pl.split != 1 || pl.comp || first < 0 || count < 1 ||
      first + count > pl.nitems_step)
    return hipErrorInvalidValue;
  const LaunchStamp ts = tl_force_stamp;
  const int2* items = pl.items_step + first;
  double* e = ew + 2 * (size_t)first;
  if (ps.periodic)
    hipExtLaunchKernelGGL((k_force_n3<true, false, 2, false, false>), dim3(count), dim3(256), 0, st,
                          ts.start, ts.stop, 0, q, posf, items, fpart, nullptr, e, pl.N, pl.Npad,
                          ps.C1);
  else
    hipExtLaunchKernelGGL((k_force_n3<false, false, 2, false, false>), dim3(count), dim3(256), 0,
                          st, ts.start, ts.stop, 0, q, posf, items, fpart, nullptr, e, pl.N, pl.Npad,
                          ps.C1);
  return hipGetLastError();
}

hipError_t launch_force_n3(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                           const uint4* q, const float4* posf, float4* fpart, float* apart,
                           bool ties, double* ew)
{
  // ew != null: the time step's variant (force partials + per-workgroup totals)
  if (pl.split == 2 && pl.n3 != 2) return hipErrorInvalidValue;
  if (pl.comp) {
    if (pl.n3 != 2 || pl.split != 1 || pl.nitems_step != pl.nitems) return hipErrorInvalidValue;
    return ew ? launch_n3_c<false>(st, pl, ps, q, posf, fpart, apart, ew, false)
              : launch_n3_c<true>(st, pl, ps, q, posf, fpart, apart, nullptr, ties);
  }
  if (ew) {
    if (pl.split == 2) return launch_n3_w<2, false, true>(st, pl, ps, q, posf, fpart, apart, ew, false);
    if (pl.n3 == 2) return launch_n3_w<2, false, false>(st, pl, ps, q, posf, fpart, apart, ew, false);
    return launch_n3_w<4, false, false>(st, pl, ps, q, posf, fpart, apart, ew, false);
  }
  if (pl.split == 2)
    return launch_n3_w<2, true, true>(st, pl, ps, q, posf, fpart, apart, nullptr, ties);
  if (pl.n3 == 2) return launch_n3_w<2, true, false>(st, pl, ps, q, posf, fpart, apart, nullptr, ties);
  return launch_n3_w<4, true, false>(st, pl, ps, q, posf, fpart, apart, nullptr, ties);
}

}  // namespace lj
