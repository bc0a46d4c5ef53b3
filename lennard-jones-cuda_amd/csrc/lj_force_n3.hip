// lj_force_n3.hip -- the force pass with every unordered pair evaluated once
// (Newton's third law), for the FP32 mode of single boxes.
//
// Physics as the reference's pair loop (MDSystem.cu:21-138, MDSystem.cpp:
// 260-309); only the order of evaluation differs.
//
// Decomposition.  The particles form groups of 512.  A work item is a pair of
// groups (g, h), h >= g, run by one workgroup of 16 waves laid out as a 4 x 4
// grid: wave (wi, wj) takes the 128 particles 512 g + 128 wi + [0, 128) as i
// (two per lane) and the 128 particles 512 h + 128 wj + [0, 128) as j.
//   * The j particles sit in LDS, stored twice in a row, so that at step t
//     lane l reads j = (l + t) mod 64 of a 64-particle sub-tile with a plain
//     immediate offset; after 64 steps every lane has met every j.
//   * i sums (force, (s/r)^6, (s/r)^6 (2 (s/r)^6 - 1)) stay in registers.
//   * j sums (minus the force, (s/r)^6 - w) go to the wave's own LDS
//     accumulators with ds_add_f32.  Within a step the 64 lanes hit 64
//     different addresses and a wave owns its accumulators, so every sum is
//     taken in a fixed order and the result is bit-reproducible.
//   * Diagonal items (g == h): waves with wi < wj evaluate their tile pair
//     once as above, waves with wi == wj evaluate both orders of their own
//     tile (i sums only, self pair excluded by index), waves with wi > wj
//     are idle.
// Output slots.  Particle p of group a receives exactly one entry per group s:
// s > a from the i side of item (a, s), s < a from the j side of item (s, a),
// s == a from the diagonal item.  The partial buffers are therefore the
// [S = groups][Npad] layout of the sliced kernel and the reduction and
// integrator kernels are shared.  Entry encoding (fx, fy, fz, bf) + af with
// e = (bf - af)/2 and W = 12 eps sum bf, as for the sliced kernel:
//   i side of a once-evaluated pair: bf = 2 b, af = b + a;
//   j side:                          bf = 0,   af = sum ((s/r)^6 - w);
//   both orders (self tile):         bf = b,   af = a.
// Per ordered pair this takes about 13.5 VALU issue slots against 22 for the
// sliced kernel (which evaluates each pair twice).
#include "lj_device.h"

namespace lj {

namespace {

using namespace dev;

constexpr int kG = 512;    // particles per group
constexpr int kTile = 128; // particles per wave tile
constexpr int kNT = 1024;  // threads per workgroup (4 x 4 waves)
constexpr int kAccPerWave = 2 * 4 * 128;  // sub-tiles x components x doubled tile

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

// one tile pair evaluated once: i sums in a1/a2, j sums into the wave's LDS
// accumulators A (4 components x 128, per sub-tile)
template <bool PERIODIC, bool TIES, bool PARTIAL>
__device__ __forceinline__ void pair_tile(const uint4* P0, float* A0, int lane,
                                          const PosT<PERIODIC>& pi1, const PosT<PERIODIC>& pi2,
                                          bool v1, bool v2, float C1, Acc& a1, Acc& a2)
{
  for (int st = 0; st < 2; ++st) {
    const uint4* P = P0 + st * 128;
    float* AX = A0 + st * 512;
    float* AY = AX + 128;
    float* AZ = AX + 256;
    float* AM = AX + 384;
#pragma unroll 8
    for (int t = 0; t < 64; ++t) {
      const int jx = lane + t;
      const uint4 pj = P[jx];
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
      const float cx = __builtin_fmaf(-dx2, s2, -(dx1 * s1));
      const float cy = __builtin_fmaf(-dy2, s2, -(dy1 * s1));
      const float cz = __builtin_fmaf(-dz2, s2, -(dz1 * s1));
      const float cm = (e1 - w1) + (e2 - w2);
      atomicAdd(AX + jx, cx);
      atomicAdd(AY + jx, cy);
      atomicAdd(AZ + jx, cz);
      atomicAdd(AM + jx, cm);
    }
  }
}

// both orders of the wave's own tile, i sums only, self pair excluded
template <bool PERIODIC, bool TIES>
__device__ __forceinline__ void self_tile(const uint4* P0, int lane, int i1, int i2, int jbase,
                                          const PosT<PERIODIC>& pi1, const PosT<PERIODIC>& pi2,
                                          bool v1, bool v2, float C1, Acc& a1, Acc& a2)
{
  for (int st = 0; st < 2; ++st) {
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

template <bool PERIODIC, bool TIES>
__global__ __launch_bounds__(kNT) void k_force_n3(const uint4* __restrict__ q,
                                                  const float4* __restrict__ pf,
                                                  const int2* __restrict__ items,
                                                  float4* __restrict__ fpart,
                                                  float* __restrict__ apart, int N, int Npad,
                                                  float C1)
{
  __shared__ uint4 s_pos[4 * 2 * 128];          // [wj][sub-tile][doubled 64]
  __shared__ float s_acc[16 * kAccPerWave];     // [wave][sub-tile][component][doubled 64]
  const int2 it = items[blockIdx.x];
  const int g = it.x, h = it.y;
  const int T = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(T >> 6);
  const int lane = T & 63;
  const int wi = wave >> 2, wj = wave & 3;
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
  for (int k = T; k < 16 * kAccPerWave / 4; k += kNT)
    reinterpret_cast<float4*>(s_acc)[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();

  const int cls = diag ? (wi < wj ? 1 : (wi == wj ? 2 : 0)) : 1;  // pair tile / self tile / idle
  const int i1 = kG * g + kTile * wi + lane, i2 = i1 + 64;
  const bool v1 = i1 < N, v2 = i2 < N;
  const PosT<PERIODIC> pi1 = load_pos<PERIODIC>(q, pf, v1 ? i1 : N - 1);
  const PosT<PERIODIC> pi2 = load_pos<PERIODIC>(q, pf, v2 ? i2 : N - 1);
  Acc a1, a2;
  acc_zero(a1);
  acc_zero(a2);
  const uint4* P0 = s_pos + wj * 256;
  if (cls == 1) {
    float* A0 = s_acc + wave * kAccPerWave;
    if (partial)
      pair_tile<PERIODIC, TIES, true>(P0, A0, lane, pi1, pi2, v1, v2, C1, a1, a2);
    else
      pair_tile<PERIODIC, TIES, false>(P0, A0, lane, pi1, pi2, v1, v2, C1, a1, a2);
  } else if (cls == 2) {
    self_tile<PERIODIC, TIES>(P0, lane, i1, i2, kG * h + kTile * wj, pi1, pi2, v1, v2, C1, a1, a2);
  }
  __syncthreads();

  // j sums of particle 512 h + T over the four i rows, in row order
  float jf[4] = {0.f, 0.f, 0.f, 0.f};
  if (T < kG) {
    const int col = T >> 7, st = (T >> 6) & 1, idx = T & 63;
    for (int r = 0; r < 4; ++r) {
      const float* A = s_acc + (r * 4 + col) * kAccPerWave + st * 512;
#pragma unroll
      for (int c = 0; c < 4; ++c) jf[c] += A[c * 128 + idx] + A[c * 128 + idx + 64];
    }
  }
  __syncthreads();

  // i sums through LDS: ib[wj][component][512]
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
    for (int c2 = 0; c2 < 4; ++c2)
#pragma unroll
      for (int c = 0; c < 5; ++c) v[c] += ib[(c2 * 5 + c) * kG + T];
    if (diag) {
      v[0] += jf[0];
      v[1] += jf[1];
      v[2] += jf[2];
      v[4] += jf[3];
    }
    const int i = kG * g + T;
    if (i < N) {
      fpart[(size_t)h * Npad + i] = make_float4(v[0], v[1], v[2], v[3]);
      apart[(size_t)h * Npad + i] = v[4];
    }
    if (!diag) {
      const int j = kG * h + T;
      if (j < N) {
        fpart[(size_t)g * Npad + j] = make_float4(jf[0], jf[1], jf[2], 0.f);
        apart[(size_t)g * Npad + j] = jf[3];
      }
    }
  }
}

}  // namespace

ForcePlan make_n3_plan(int N)
{
  ForcePlan pl;
  pl.N = N;
  pl.n3 = 1;
  pl.ng = (N + kG - 1) / kG;
  pl.S = pl.ng;
  pl.jchunk = kG;
  pl.Npad = (N + 63) & ~63;
  pl.nitems = pl.ng * (pl.ng + 1) / 2;
  return pl;
}

bool n3_eligible(int N) { return (N + kG - 1) / kG >= 24; }

void n3_items(int ng, int2* out)
{
  // off-diagonal items first, the lighter diagonal ones last
  int k = 0;
  for (int g = 0; g < ng; ++g)
    for (int h = g + 1; h < ng; ++h) out[k++] = make_int2(g, h);
  for (int g = 0; g < ng; ++g) out[k++] = make_int2(g, g);
}

hipError_t launch_force_n3(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                           const uint4* q, const float4* posf, float4* fpart, float* apart,
                           bool ties)
{
  dim3 grid(pl.nitems), block(kNT);
  if (ps.periodic) {
    if (ties)
      hipLaunchKernelGGL((k_force_n3<true, true>), grid, block, 0, st, q, posf, pl.items, fpart,
                         apart, pl.N, pl.Npad, ps.C1);
    else
      hipLaunchKernelGGL((k_force_n3<true, false>), grid, block, 0, st, q, posf, pl.items, fpart,
                         apart, pl.N, pl.Npad, ps.C1);
  } else {
    hipLaunchKernelGGL((k_force_n3<false, false>), grid, block, 0, st, q, posf, pl.items, fpart,
                       apart, pl.N, pl.Npad, ps.C1);
  }
  return hipGetLastError();
}

}  // namespace lj
