// lj_force_n3.hip -- the force pass with every unordered pair evaluated once
// (Newton's third law), for the FP32 mode of single boxes.
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
// The time-step variant's loop is 24 VALU instructions per unordered pair
// (12.4 per ordered pair over the whole kernel, measured) against 21.5 per
// ordered pair for the sliced kernel, which evaluates each pair twice.
#include <cstdlib>

#include "lj_device.h"

namespace lj {

namespace {

using namespace dev;

constexpr int kTile = 128;  // particles per wave tile
constexpr int kAccPerWave = 2 * 4 * 64;  // sub-tiles x components x tile

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
// nsub: sub-tiles holding a valid j (1 when the second one is past N)
template <bool PERIODIC, bool TIES, bool PARTIAL, bool ENERGY>
__device__ __forceinline__ void pair_tile(const uint4* P0, int lane, const PosT<PERIODIC>& pi1,
                                          const PosT<PERIODIC>& pi2, bool v1, bool v2, float C1,
                                          int nsub, Acc& a1, Acc& a2, float (&J)[2][4])
{
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    if (st >= nsub) break;  // wave-uniform
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
                                          bool v1, bool v2, float C1, int nsub, Acc& a1, Acc& a2)
{
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    if (st >= nsub) break;  // wave-uniform
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

// W x W waves per workgroup, groups of G = 128 W particles.  Eight waves per
// SIMD (at most 64 VGPRs; the loop body is unchanged by the limit, no spills).
// ENERGY = false (the time step): no per-particle energies; the partial
// entries carry the force only (w = 0, no apart), and each workgroup writes its
// totals ew[2 item] = sum over ordered pairs of ((s/r)^12 - (s/r)^6) and
// ew[2 item + 1] = sum of w = (s/r)^6 (2 (s/r)^6 - 1), the finaliser's E and W
// terms.  The j side then carries three sums instead of four.
// One group pair (g, h) of one box: q/pf/fpart/apart point at the box's
// arrays, ew (ENERGY = false) at the item's two totals.
template <bool PERIODIC, bool TIES, int W, bool ENERGY>
__device__ __forceinline__ void n3_item(const uint4* __restrict__ q, const float4* __restrict__ pf,
                                        float4* __restrict__ fpart, float* __restrict__ apart,
                                        double* __restrict__ ew, int N, int Npad, float C1, int g,
                                        int h)
{
  constexpr int kG = kTile * W;
  constexpr int kAcc = W * W * kAccPerWave > W * 5 * kG ? W * W * kAccPerWave : W * 5 * kG;
  __shared__ uint4 s_pos[W * 2 * 128];  // [wj][sub-tile][doubled 64]
  // j sums [wave][sub-tile][component][64], then i sums [wj][component][G]
  __shared__ float s_acc[kAcc];
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
  const int nsub = jt0 + 64 < N ? 2 : 1;
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
  if (cls == 1) {
    if (partial)
      pair_tile<PERIODIC, TIES, true, ENERGY>(P0, lane, pi1, pi2, v1, v2, C1, nsub, a1, a2, J);
    else
      pair_tile<PERIODIC, TIES, false, ENERGY>(P0, lane, pi1, pi2, v1, v2, C1, 2, a1, a2, J);
  } else if (cls == 2) {
    self_tile<PERIODIC, TIES>(P0, lane, i1, i2, kG * h + kTile * wj, pi1, pi2, v1, v2, C1, nsub, a1,
                              a2);
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
    __shared__ double s_ew[W * W][2];
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
    if (i < N) {
      store_part(fpart + (size_t)h * Npad + i, v[0], v[1], v[2], ENERGY ? v[3] : 0.f);
      if constexpr (ENERGY) store_part(apart + (size_t)h * Npad + i, v[4]);
    }
    if (!diag) {
      const int j = kG * h + T;
      if (j < N) {
        store_part(fpart + (size_t)g * Npad + j, jf[0], jf[1], jf[2], 0.f);
        if constexpr (ENERGY) store_part(apart + (size_t)g * Npad + j, jf[3]);
      }
    }
  }
}

// FIN: one workgroup besides the items computes the observables of the
// previous step from its block rows and force totals (`fin`: that step's
// arguments, its own ew buffer), in place of a separate finaliser launch
// between the steps.  It is block 0, so that it is dispatched with the first
// wave of workgroups: as the last block it would wait for a free slot and
// finish after the items.
template <bool PERIODIC, bool TIES, int W, bool ENERGY, bool FIN>
__global__ __launch_bounds__(64 * W * W) __attribute__((amdgpu_waves_per_eu(8))) void k_force_n3(
    const uint4* __restrict__ q, const float4* __restrict__ pf, const int2* __restrict__ items,
    float4* __restrict__ fpart, float* __restrict__ apart, double* __restrict__ ew, int N, int Npad,
    float C1, StepArgs fin)
{
  if constexpr (FIN) {
    if (blockIdx.x == 0) {
      step_observables<W * W>(fin, fin.nrow, true, true);
      return;
    }
  }
  const int k = FIN ? (int)blockIdx.x - 1 : (int)blockIdx.x;  // item
  const int2 it = items[k];
  n3_item<PERIODIC, TIES, W, ENERGY>(q, pf, fpart, apart, ENERGY ? nullptr : ew + 2 * (size_t)k, N,
                                     Npad, C1, it.x, it.y);
}

// batched ensemble: item = (box, g, h) over all boxes; each box's partials at
// its own rows (pbase, S = groups of 256, Npad); ENERGY: per-particle energies
// kept (single evaluations), else the time step's variant with the item's
// (E, W) totals at ew[2 e]
template <bool PERIODIC, bool TIES, bool ENERGY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_batch_force_n3(
    const uint4* __restrict__ q, const float4* __restrict__ pf, float4* __restrict__ fpart,
    float* __restrict__ apart, const BoxParam* __restrict__ box, const BatchItem* __restrict__ items,
    double* __restrict__ ew)
{
  const BatchItem w = items[blockIdx.x];
  const BoxParam bp = box[w.box];
  n3_item<PERIODIC, TIES, 2, ENERGY>(q + bp.off, pf + bp.off, fpart + bp.pbase, apart + bp.pbase,
                                     ENERGY ? nullptr : ew + 2 * (size_t)w.e, bp.N, bp.Npad, bp.C1,
                                     w.i0, w.s);
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

ForcePlan make_n3_plan(int N)
{
  ForcePlan pl;
  const int W = n3_waves(N);
  const int G = kTile * W;
  pl.N = N;
  pl.n3 = W;
  pl.ng = (N + G - 1) / G;
  pl.S = pl.ng;
  pl.jchunk = G;
  pl.Npad = (N + 63) & ~63;
  pl.nitems = pl.ng * (pl.ng + 1) / 2;
  return pl;
}

bool n3_eligible(int N)
{
  // LJ_PAIR_ONCE=0 / 1 turns the kernel off / on at any size (tuning)
  const char* e = std::getenv("LJ_PAIR_ONCE");
  if (e && (e[0] == '0' || e[0] == '1')) return e[0] == '1' && N >= 256;
  // measured against the sliced kernel (time step, TVN, rho* = 0.85): N = 6144
  // 33.6 vs 37.3 us per step (force pass 22.4 vs 22.9 us), N = 4096 28.7 vs
  // 26.1 us (few group pairs leave most CUs idle: 18.6 vs 12.0 us)
  return N >= 6144;
}

void n3_items(int ng, int2* out)
{
  // off-diagonal items first, the lighter diagonal ones last
  int k = 0;
  for (int g = 0; g < ng; ++g)
    for (int h = g + 1; h < ng; ++h) out[k++] = make_int2(g, h);
  for (int g = 0; g < ng; ++g) out[k++] = make_int2(g, g);
}

namespace {

template <int W, bool ENERGY, bool FIN>
void launch_n3_f(hipStream_t st, const ForcePlan& pl, const PairScalars& ps, const uint4* q,
                 const float4* posf, float4* fpart, float* apart, double* ew, bool ties,
                 const StepArgs& fin)
{
  dim3 grid(pl.nitems + (FIN ? 1 : 0)), block(64 * W * W);
  if (ps.periodic) {
    if (ties)
      hipLaunchKernelGGL((k_force_n3<true, true, W, ENERGY, FIN>), grid, block, 0, st, q, posf,
                         pl.items, fpart, apart, ew, pl.N, pl.Npad, ps.C1, fin);
    else
      hipLaunchKernelGGL((k_force_n3<true, false, W, ENERGY, FIN>), grid, block, 0, st, q, posf,
                         pl.items, fpart, apart, ew, pl.N, pl.Npad, ps.C1, fin);
  } else {
    hipLaunchKernelGGL((k_force_n3<false, false, W, ENERGY, FIN>), grid, block, 0, st, q, posf,
                       pl.items, fpart, apart, ew, pl.N, pl.Npad, ps.C1, fin);
  }
}

template <int W, bool ENERGY>
hipError_t launch_n3_w(hipStream_t st, const ForcePlan& pl, const PairScalars& ps, const uint4* q,
                       const float4* posf, float4* fpart, float* apart, double* ew, bool ties,
                       const StepArgs* fin)
{
  if constexpr (!ENERGY) {  // the finaliser workgroup rides on the time step's variant only
    if (fin) {
      launch_n3_f<W, ENERGY, true>(st, pl, ps, q, posf, fpart, apart, ew, ties, *fin);
      return hipGetLastError();
    }
  }
  if (fin) return hipErrorInvalidValue;
  launch_n3_f<W, ENERGY, false>(st, pl, ps, q, posf, fpart, apart, ew, ties, StepArgs{});
  return hipGetLastError();
}

}  // namespace

hipError_t launch_batch_force_n3(hipStream_t st, const BatchArgs& a, bool periodic, bool ties)
{
  dim3 grid(a.nitems), block(256);
  if (a.lean) {  // the time step (no exact-tie rule there)
    if (ties || !a.ew) return hipErrorInvalidValue;
    if (periodic)
      hipLaunchKernelGGL((k_batch_force_n3<true, false, false>), grid, block, 0, st, a.q, a.pos,
                         a.fpart, a.apart, a.box, a.items, a.ew);
    else
      hipLaunchKernelGGL((k_batch_force_n3<false, false, false>), grid, block, 0, st, a.q, a.pos,
                         a.fpart, a.apart, a.box, a.items, a.ew);
    return hipGetLastError();
  }
  if (periodic) {
    if (ties)
      hipLaunchKernelGGL((k_batch_force_n3<true, true, true>), grid, block, 0, st, a.q, a.pos,
                         a.fpart, a.apart, a.box, a.items, a.ew);
    else
      hipLaunchKernelGGL((k_batch_force_n3<true, false, true>), grid, block, 0, st, a.q, a.pos,
                         a.fpart, a.apart, a.box, a.items, a.ew);
  } else {
    hipLaunchKernelGGL((k_batch_force_n3<false, false, true>), grid, block, 0, st, a.q, a.pos,
                       a.fpart, a.apart, a.box, a.items, a.ew);
  }
  return hipGetLastError();
}

hipError_t launch_force_n3(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                           const uint4* q, const float4* posf, float4* fpart, float* apart,
                           bool ties, double* ew, const StepArgs* fin)
{
  // ew != null: the time step's variant (force partials + per-workgroup totals),
  // with the previous step's observables when fin != null
  if (ew) {
    if (pl.n3 == 2) return launch_n3_w<2, false>(st, pl, ps, q, posf, fpart, apart, ew, false, fin);
    return launch_n3_w<4, false>(st, pl, ps, q, posf, fpart, apart, ew, false, fin);
  }
  if (pl.n3 == 2) return launch_n3_w<2, true>(st, pl, ps, q, posf, fpart, apart, nullptr, ties, nullptr);
  return launch_n3_w<4, true>(st, pl, ps, q, posf, fpart, apart, nullptr, ties, nullptr);
}

}  // namespace lj
