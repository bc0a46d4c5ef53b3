// lj_batch.hip -- many independent boxes of uneven N, density and temperature
// advanced together: one persistent force launch drains a work queue of
// (box, i-tile, j-slice) items over all boxes; the per-particle kernels run
// over workgroups that never straddle two boxes, so every per-box reduction
// stays a fixed-order FP64 sum.  Physics per box is the single-box step of
// lj_kernels.hip (reference MDSystem.cpp:438-583, MDSystem.cu:21-138).
#include "lj_device.h"

namespace lj {

namespace {

using namespace dev;

template <bool PERIODIC, int IPL, bool TIES>
__device__ __forceinline__ void batch_item(const uint4* __restrict__ q, const float4* __restrict__ pf,
                                           float4* __restrict__ fpart, float* __restrict__ apart,
                                           const BoxParam& bp, int i0, int s, int lane)
{
  const int N = bp.N;
  const int j0 = s * bp.jchunk;
  const int j1 = min(N, j0 + bp.jchunk);
  const uint4* qb = q + bp.off;
  const float4* pb = pf + bp.off;
  PosT<PERIODIC> pi[IPL];
  int ii[IPL];
  Acc hi[IPL], lo[IPL];
#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    ii[k] = i0 + k * 64 + lane;
    pi[k] = load_pos<PERIODIC>(qb, pb, min(ii[k], N - 1));
    acc_zero(hi[k]);
    acc_zero(lo[k]);
  }
  const int d0 = max(j0, i0), d1 = min(j1, i0 + 64 * IPL);
  if (d0 >= d1) {
    j_range<PERIODIC, 0, IPL, false, TIES>(qb, pb, j0, j1, pi, ii, bp.C1, 0.f, hi, lo);
  } else {
    j_range<PERIODIC, 0, IPL, false, TIES>(qb, pb, j0, d0, pi, ii, bp.C1, 0.f, hi, lo);
    j_range<PERIODIC, 1, IPL, false, TIES>(qb, pb, d0, d1, pi, ii, bp.C1, 0.f, hi, lo);
    j_range<PERIODIC, 0, IPL, false, TIES>(qb, pb, d1, j1, pi, ii, bp.C1, 0.f, hi, lo);
  }
#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    if (ii[k] < N) {
      const size_t o = (size_t)bp.pbase + (size_t)s * bp.Npad + ii[k];
      store_part(fpart + o, hi[k].fx, hi[k].fy, hi[k].fz, hi[k].b);
      store_part(apart + o, hi[k].a);
    }
  }
}

// PERSIST: waves pull items from the queue head until it passes nitems (every
// wave reaches that exit); otherwise wave w of the grid takes item w.
template <bool PERIODIC, int IPL, bool PERSIST, bool TIES>
__global__ __launch_bounds__(kBlock) void k_batch_force(const uint4* __restrict__ q,
                                                        const float4* __restrict__ pf,
                                                        float4* __restrict__ fpart,
                                                        float* __restrict__ apart,
                                                        const BoxParam* __restrict__ box,
                                                        const BatchItem* __restrict__ items,
                                                        unsigned* __restrict__ counter, int nitems)
{
  const int lane = threadIdx.x & 63;
  if constexpr (PERSIST) {
    for (;;) {
      unsigned it = 0;
      if (lane == 0) it = atomicAdd(counter, 1u);
      it = __builtin_amdgcn_readfirstlane(__shfl(it, 0, 64));
      if (it >= (unsigned)nitems) break;
      const BatchItem w = items[it];
      const BoxParam bp = box[w.box];
      batch_item<PERIODIC, IPL, TIES>(q, pf, fpart, apart, bp, w.i0, w.s, lane);
    }
  } else {
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int it = blockIdx.x * kWaves + wv;
    if (it >= nitems) return;
    const BatchItem w = items[it];
    const BoxParam bp = box[w.box];
    batch_item<PERIODIC, IPL, TIES>(q, pf, fpart, apart, bp, w.i0, w.s, lane);
  }
}

struct PIdx {
  int b, i, gi;
  bool valid;
};

__device__ __forceinline__ PIdx pidx(const BatchArgs& a)
{
  const int2 pb = a.pblock[blockIdx.x];
  PIdx r;
  r.b = pb.x;
  r.i = pb.y + (int)threadIdx.x;
  const BoxParam& bp = a.box[r.b];
  r.valid = r.i < bp.N;
  r.gi = bp.off + r.i;
  return r;
}

__global__ __launch_bounds__(kBlock) void k_batch_prep_q(BatchArgs a)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0u;
  const PIdx p = pidx(a);
  if (!p.valid) return;
  const double qs = a.box[p.b].qscale;
  const float4 x = a.pos[p.gi];
  a.q[p.gi] = make_uint4(fixq(x.x, qs), fixq(x.y, qs), fixq(x.z, qs), 0u);
}

template <int CANON>
__global__ __launch_bounds__(kBlock) void k_batch_drift(BatchArgs a)
{
#pragma clang fp contract(off)
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0u;  // queue of this step's force pass
  const PIdx pp = pidx(a);
  if (!pp.valid) return;
  const int i = pp.gi;
  float4 p = a.pos[i];
  float4 v = a.vel[i];
  float4 tf;
  drift_one<CANON>(p, v, tf, a.force[i], a.dt);  // MDSystem.cpp:447-453, 477-479
  if constexpr (CANON == LJ_EVN) a.vel[i] = v;
  else a.tF[i] = tf;
  a.pos[i] = p;
  if (a.bc == LJ_BC_PERIODIC) {
    const double qs = a.box[pp.b].qscale;
    a.q[i] = make_uint4(fixq(p.x, qs), fixq(p.y, qs), fixq(p.z, qs), 0u);
  }
}

// the next step's drift of one particle at the end of a kick kernel (the
// arithmetic of k_batch_drift on the values the kick just computed); block 0
// also resets the work-queue head for the next force pass
template <int CANON>
__device__ __forceinline__ void batch_drift_next(const BatchArgs& a, const PIdx& pp, float4 p,
                                                 float4 v, const float4& f)
{
  float4 tf;
  drift_one<CANON>(p, v, tf, f, a.dt);
  if constexpr (CANON == LJ_EVN) a.vel[pp.gi] = v;
  else a.tF[pp.gi] = tf;
  a.pos[pp.gi] = p;
  if (a.bc == LJ_BC_PERIODIC) {
    const double qs = a.box[pp.b].qscale;
    a.q[pp.gi] = make_uint4(fixq(p.x, qs), fixq(p.y, qs), fixq(p.z, qs), 0u);
  }
}

__device__ __forceinline__ Partial box_slices(const BatchArgs& a, const BoxParam& bp, int i)
{
  if (a.lean)  // force-only partials of the time step's pair-once pass (a = b = 0)
    return sum_slices<false>(a.fpart + bp.pbase, a.apart + bp.pbase, i, bp.S, bp.Npad);
  return sum_slices<true>(a.fpart + bp.pbase, a.apart + bp.pbase, i, bp.S, bp.Npad);
}

__device__ __forceinline__ float inside(const float4& p, const BoxParam& bp)
{
  return (p.x >= bp.sub_lo && p.x <= bp.sub_hi && p.y >= bp.sub_lo && p.y <= bp.sub_hi &&
          p.z >= bp.sub_lo && p.z <= bp.sub_hi)
             ? 1.f
             : 0.f;
}

template <bool DRIFT>
__global__ __launch_bounds__(kBlock) void k_batch_kick_evn(BatchArgs a)
{
#pragma clang fp contract(off)
  if (DRIFT && blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0u;  // next step's force pass
  const PIdx pp = pidx(a);
  const BoxParam& bp = a.box[pp.b];
  double v4[4] = {0., 0., 0., 0.};
  if (pp.valid) {
    const int i = pp.gi;
    // the particle's own loads go out ahead of the slot sums (their latencies
    // overlap; the force store below would otherwise keep them after it)
    float4 v = a.vel[i];
    float4 p = a.pos[i];
    const Partial s = box_slices(a, bp, pp.i);
    const double e = 0.5 * (s.b - s.a);
    const float fx = (float)(bp.FS * s.fx), fy = (float)(bp.FS * s.fy), fz = (float)(bp.FS * s.fz);
    a.force[i] = make_float4(fx, fy, fz, (float)e);
    const double dt = a.dt;
    v.x = (float)(v.x + dt * fx / 2.);  // MDSystem.cpp:460-462
    v.y = (float)(v.y + dt * fy / 2.);
    v.z = (float)(v.z + dt * fz / 2.);
    if (a.bc != LJ_BC_NONE) {
      apply_bc_one(p, v, a.bc, bp.L);
      if (!DRIFT && a.bc == LJ_BC_PERIODIC) a.pos[i] = p;
    }
    v4[0] = kin2(v);
    v4[1] = e;
    v4[2] = s.b;
    v4[3] = inside(p, bp);
    if constexpr (DRIFT) batch_drift_next<LJ_EVN>(a, pp, p, v, make_float4(fx, fy, fz, 0.f));
    else a.vel[i] = v;
  }
  const int fld[4] = {BLK_K2, BLK_E, BLK_B, BLK_NSUB};
  block_sum_store<4>(v4, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

__global__ __launch_bounds__(kBlock) void k_batch_tvn_a(BatchArgs a)
{
#pragma clang fp contract(off)
  const PIdx pp = pidx(a);
  const BoxParam& bp = a.box[pp.b];
  double v3[3] = {0., 0., 0.};
  if (pp.valid) {
    const int i = pp.gi;
    // the particle's own loads go out ahead of the slot sums
    float4 tf = a.tF[i];
    const float4 v = a.vel[i];
    const Partial s = box_slices(a, bp, pp.i);
    const double e = 0.5 * (s.b - s.a);
    const float fx = (float)(bp.FS * s.fx), fy = (float)(bp.FS * s.fy), fz = (float)(bp.FS * s.fz);
    a.force[i] = make_float4(fx, fy, fz, (float)e);
    tf.x += 0.5f * fx;  // MDSystem.cpp:486-488
    tf.y += 0.5f * fy;
    tf.z += 0.5f * fz;
    a.tF[i] = tf;
    const double dt = a.dt;
    float4 tv;
    tv.x = (float)(v.x + dt * tf.x / 2.);  // MDSystem.cpp:493-495
    tv.y = (float)(v.y + dt * tf.y / 2.);
    tv.z = (float)(v.z + dt * tf.z / 2.);
    tv.w = 0.f;
    v3[0] = kin2(tv);
    v3[1] = e;
    v3[2] = s.b;
  }
  const int fld[3] = {BLK_TV2, BLK_E, BLK_B};
  block_sum_store<3>(v3, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

template <bool DRIFT>
__global__ __launch_bounds__(kBlock) void k_batch_tvn_b(BatchArgs a)
{
#pragma clang fp contract(off)
  if (DRIFT && blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0u;  // next step's force pass
  const PIdx pp = pidx(a);
  const BoxParam& bp = a.box[pp.b];
  const double tv2 = blk_total(a.blk + (size_t)bp.blk0 * kBlkFields, bp.nblk, BLK_TV2);
  const double Tkin = tv2 * (1. / 3. / bp.N);  // MDSystem.cpp:370
  const double chi = sqrt(bp.T0 / Tkin);      // MDSystem.cpp:499
  double v2[2] = {0., 0.};
  if (pp.valid) {
    const int i = pp.gi;
    float4 v = a.vel[i];
    const float4 tf = a.tF[i];
    const double dt = a.dt;
    v.x = (float)((2. * chi - 1.) * v.x + chi * dt * tf.x);  // MDSystem.cpp:503-505
    v.y = (float)((2. * chi - 1.) * v.y + chi * dt * tf.y);
    v.z = (float)((2. * chi - 1.) * v.z + chi * dt * tf.z);
    float4 p = a.pos[i];
    if (a.bc != LJ_BC_NONE) {
      apply_bc_one(p, v, a.bc, bp.L);
      if (!DRIFT && a.bc == LJ_BC_PERIODIC) a.pos[i] = p;
    }
    a.vel[i] = v;
    v2[0] = kin2(v);
    v2[1] = inside(p, bp);
    if constexpr (DRIFT) batch_drift_next<LJ_TVN>(a, pp, p, v, a.force[i]);
  }
  __syncthreads();
  const int fld[2] = {BLK_K2, BLK_NSUB};
  block_sum_store<2>(v2, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

__global__ __launch_bounds__(kBlock) void k_batch_reduce_forces(BatchArgs a)
{
  const PIdx pp = pidx(a);
  const BoxParam& bp = a.box[pp.b];
  double v2[2] = {0., 0.};
  if (pp.valid) {
    const Partial s = box_slices(a, bp, pp.i);
    const double e = 0.5 * (s.b - s.a);
    a.force[pp.gi] = make_float4((float)(bp.FS * s.fx), (float)(bp.FS * s.fy),
                                 (float)(bp.FS * s.fz), (float)e);
    v2[0] = e;
    v2[1] = s.b;
  }
  const int fld[2] = {BLK_E, BLK_B};
  block_sum_store<2>(v2, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

__global__ __launch_bounds__(kBlock) void k_batch_kinetic(BatchArgs a)
{
  const PIdx pp = pidx(a);
  const BoxParam& bp = a.box[pp.b];
  double v2[2] = {0., 0.};
  if (pp.valid) {
    v2[0] = kin2(a.vel[pp.gi]);
    v2[1] = inside(a.pos[pp.gi], bp);
  }
  const int fld[2] = {BLK_K2, BLK_NSUB};
  block_sum_store<2>(v2, a.blk + (size_t)blockIdx.x * kBlkFields, fld);
}

// one workgroup per box: observables (MDSystem.cpp:325-359) and, when
// advancing, the accumulated fluctuation statistics of the box
__global__ __launch_bounds__(kBlock) void k_batch_finalize(BatchArgs a, int advance)
{
  const int b = blockIdx.x;
  const BoxParam& bp = a.box[b];
  double v[4] = {0., 0., 0., 0.};
  const double* rows = a.blk + (size_t)bp.blk0 * kBlkFields;
  for (int r = threadIdx.x; r < bp.nblk; r += kBlock) {
    v[0] += rows[(size_t)r * kBlkFields + BLK_K2];
    v[1] += rows[(size_t)r * kBlkFields + BLK_E];
    v[2] += rows[(size_t)r * kBlkFields + BLK_B];
    v[3] += rows[(size_t)r * kBlkFields + BLK_NSUB];
  }
  if (a.lean)  // the box's per-item totals of the step's force pass, in slot order
    for (int k = threadIdx.x; k < bp.nitem; k += kBlock) {
      v[1] += a.ew[2 * (size_t)(bp.ew0 + k)];
      v[2] += a.ew[2 * (size_t)(bp.ew0 + k) + 1];
    }
  block_allsum<4>(v);
  if (threadIdx.x != 0) return;
  lj_obs* o = a.obs + b;
  const int N = bp.N;
  const double K = v[0] / 2.;
  const double T = 2. * K / 3. / N;
  o->V = 2. * a.eps * v[1];
  o->W = 12. * a.eps * v[2];
  const double P = (o->W / 3. + N * T) / (N / bp.rho);
  o->K = K;
  o->T = T;
  o->P = P;
  o->U = K + o->V;
  o->Pshear = 0.;
  if (advance) {
    o->t += a.dt;
    const int64_t k = o->step - (a.hist_base ? a.hist_base[b] : 0);
    o->step += 1;
    if (a.hist && k >= 0 && k < a.hist_cap) {
      double* h = a.hist + 5 * ((size_t)k * a.nboxes + b);
      h[0] = o->U;
      h[1] = T;
      h[2] = K;
      h[3] = o->V;
      h[4] = P;
    }
    o->av_iters += 1;
    o->av_U_tot += o->U;
    o->av_p_tot += P;
    o->av_T_tot += T;
    double* st = a.stats + (size_t)b * kNStat;
    const double n = v[3];
    st[ST_SAMPLES] += 1.;
    st[ST_U] += o->U;
    st[ST_U2] += o->U * o->U;
    st[ST_V] += o->V;
    st[ST_V2] += o->V * o->V;
    st[ST_K] += K;
    st[ST_K2] += K * K;
    st[ST_P] += P;
    st[ST_P2] += P * P;
    st[ST_N1] += n;
    st[ST_N2] += n * n;
    st[ST_N3] += n * n * n;
    st[ST_N4] += n * n * n * n;
  }
}

__global__ void k_batch_set_hist_base(const lj_obs* obs, int64_t* base, int nboxes)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nboxes) base[i] = obs[i].step;
}

__global__ void k_batch_zero_stats(double* stats, int n, lj_obs* obs, int nboxes)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) stats[i] = 0.;
  if (i < nboxes) {
    obs[i].av_U_tot = obs[i].av_T_tot = obs[i].av_p_tot = 0.;
    obs[i].av_iters = 0;
  }
}

template <bool PERIODIC, int IPL, bool TIES>
void launch_batch_force_t(hipStream_t st, const BatchArgs& a)
{
  if (a.persistent_grid > 0)
    hipLaunchKernelGGL((k_batch_force<PERIODIC, IPL, true, TIES>), dim3(a.persistent_grid),
                       dim3(kBlock), 0, st, a.q, a.pos, a.fpart, a.apart, a.box, a.items,
                       a.counter, a.nitems);
  else
    hipLaunchKernelGGL((k_batch_force<PERIODIC, IPL, false, TIES>),
                       dim3((a.nitems + kWaves - 1) / kWaves), dim3(kBlock), 0, st, a.q, a.pos,
                       a.fpart, a.apart, a.box, a.items, a.counter, a.nitems);
}

}  // namespace

hipError_t launch_batch_prep_q(hipStream_t st, const BatchArgs& a)
{
  hipLaunchKernelGGL(k_batch_prep_q, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_batch_drift(hipStream_t st, const BatchArgs& a)
{
  if (a.canonical == LJ_EVN)
    hipLaunchKernelGGL(k_batch_drift<LJ_EVN>, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  else
    hipLaunchKernelGGL(k_batch_drift<LJ_TVN>, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_batch_force(hipStream_t st, const BatchArgs& a, bool periodic, bool ties)
{
  if (a.pair_once) return launch_batch_force_n3(st, a, periodic, ties);
  if (periodic) {
    if (ties) {
      if (a.ipl == 2) launch_batch_force_t<true, 2, true>(st, a);
      else launch_batch_force_t<true, 1, true>(st, a);
    } else {
      if (a.ipl == 2) launch_batch_force_t<true, 2, false>(st, a);
      else launch_batch_force_t<true, 1, false>(st, a);
    }
  } else {
    if (a.ipl == 2) launch_batch_force_t<false, 2, false>(st, a);
    else launch_batch_force_t<false, 1, false>(st, a);
  }
  return hipGetLastError();
}

hipError_t launch_batch_kick_evn(hipStream_t st, const BatchArgs& a, bool drift_next)
{
  if (drift_next) hipLaunchKernelGGL(k_batch_kick_evn<true>, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  else hipLaunchKernelGGL(k_batch_kick_evn<false>, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_batch_tvn_a(hipStream_t st, const BatchArgs& a)
{
  hipLaunchKernelGGL(k_batch_tvn_a, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_batch_tvn_b(hipStream_t st, const BatchArgs& a, bool drift_next)
{
  if (drift_next) hipLaunchKernelGGL(k_batch_tvn_b<true>, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  else hipLaunchKernelGGL(k_batch_tvn_b<false>, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_batch_finalize(hipStream_t st, const BatchArgs& a, int advance)
{
  hipLaunchKernelGGL(k_batch_finalize, dim3(a.nboxes), dim3(kBlock), 0, st, a, advance);
  return hipGetLastError();
}

hipError_t launch_batch_reduce_forces(hipStream_t st, const BatchArgs& a)
{
  hipLaunchKernelGGL(k_batch_reduce_forces, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_batch_kinetic(hipStream_t st, const BatchArgs& a)
{
  hipLaunchKernelGGL(k_batch_kinetic, dim3(a.nblocks), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_batch_set_hist_base(hipStream_t st, const BatchArgs& a, int64_t* hist_base)
{
  hipLaunchKernelGGL(k_batch_set_hist_base, dim3((a.nboxes + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     st, a.obs, hist_base, a.nboxes);
  return hipGetLastError();
}

hipError_t launch_batch_zero_stats(hipStream_t st, const BatchArgs& a)
{
  const int n = a.nboxes * kNStat;
  const int m = n > a.nboxes ? n : a.nboxes;
  hipLaunchKernelGGL(k_batch_zero_stats, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     a.stats, n, a.obs, a.nboxes);
  return hipGetLastError();
}

}  // namespace lj
