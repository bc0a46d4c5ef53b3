// lj_kernels.hip -- gfx950 kernels of the Lennard-Jones time step.
//
// Force pass (the hot path; reference MDSystem.cu:21-138, MDSystem.cpp:260-309):
//   * one i particle per lane (or IPL per lane), wave64, 4 waves per workgroup;
//   * the j range is split into S slices (grid.y) so that N = 16k yields
//     8 waves per SIMD on 256 CUs; each wave streams its slice's j particles
//     through SGPRs (uniform scalar loads; every VALU op takes the j operand
//     straight from an SGPR, so there is no LDS traffic and no barrier);
//   * periodic boxes use 32-bit fixed-point coordinates q = x * 2^32 / L:
//     the unsigned difference wraps modulo 2^32, which *is* the minimum image
//     (no round/div per pair); the wall / open box uses the float coordinates;
//   * per pair: 3 sub + 3 cvt + 3 (r^2) + v_rcp_f32 + 1 (scale) + 2 ((s/r)^6)
//     + 3 (w, s) + 3 (force) + 2 (energy sums) VALU ops;
//   * partial sums per (slice, i) are written once; the cross-slice sum is a
//     fixed-order FP64 pass fused into the integrator kernels.
//   Energy and virial come from two per-i sums: a = sum (s/r)^6 and
//   b = sum (s/r)^6 (2 (s/r)^6 - 1):  e_i = sum ((s/r)^12 - (s/r)^6) = (b - a)/2,
//   per-i virial sum_j r.f_ij = 24 eps b.
// RDF pass (MDSystem.cu:39-42, MDSystem.cpp:281-285): LDS-privatised uint32
// histogram per workgroup, flushed with one 64-bit atomic per bin.
// Integrator (MDSystem.cpp:438-510), BCs (:406-436), observables (:325-359):
// per-particle kernels with the reference's double/float rounding pattern and
// deterministic FP64 block reductions.
#include <cstdlib>

#include "lj_device.h"
#include "lj_rdf_device.h"

namespace lj {

namespace {

using namespace dev;
using namespace rdfdev;

// grid: (gx, S); block: kBlock.  fpart[s*Npad + i] = (fx, fy, fz, b), apart = a,
// in internal units (force = FS * f), for the i rows [ibase, iend): all N, or
// one GPU's share of a pass sharded by i (lj_forces_shard).
template <bool PERIODIC, bool GUARDALL, int IPL, bool COMP, bool TIES>
__global__ __launch_bounds__(kBlock) void k_force(const uint4* __restrict__ q,
                                                  const float4* __restrict__ pf,
                                                  float4* __restrict__ fpart,
                                                  float* __restrict__ apart, int N, int Npad,
                                                  int jchunk, float C1, float thr, int ibase,
                                                  int iend)
{
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = ibase + (blockIdx.x * kWaves + wv) * (64 * IPL);
  if (i0 >= iend) return;
  const int s = blockIdx.y;
  const int j0 = s * jchunk;
  const int j1 = min(N, j0 + jchunk);

  PosT<PERIODIC> pi[IPL];
  int ii[IPL];
  Acc hi[IPL], lo[IPL];
#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    ii[k] = i0 + k * 64 + lane;
    pi[k] = load_pos<PERIODIC>(q, pf, min(ii[k], iend - 1));
    acc_zero(hi[k]);
    acc_zero(lo[k]);
  }

  if constexpr (GUARDALL) {
    j_range<PERIODIC, 2, IPL, COMP, TIES>(q, pf, j0, j1, pi, ii, C1, thr, hi, lo);
  } else {
    // the self pair lies only in [i0, i0 + 64*IPL): guard that segment alone
    const int d0 = max(j0, i0), d1 = min(j1, i0 + 64 * IPL);
    if (d0 >= d1) {
      j_range<PERIODIC, 0, IPL, COMP, TIES>(q, pf, j0, j1, pi, ii, C1, thr, hi, lo);
    } else {
      j_range<PERIODIC, 0, IPL, COMP, TIES>(q, pf, j0, d0, pi, ii, C1, thr, hi, lo);
      j_range<PERIODIC, 1, IPL, COMP, TIES>(q, pf, d0, d1, pi, ii, C1, thr, hi, lo);
      j_range<PERIODIC, 0, IPL, COMP, TIES>(q, pf, d1, j1, pi, ii, C1, thr, hi, lo);
    }
  }

#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    if (ii[k] < iend) {
      const size_t o = (size_t)s * Npad + ii[k];
      if constexpr (COMP) {
        store_part(fpart + o, hi[k].fx + lo[k].fx, hi[k].fy + lo[k].fy, hi[k].fz + lo[k].fz,
                   hi[k].b + lo[k].b);
        store_part(apart + o, hi[k].a + lo[k].a);
      } else {
        store_part(fpart + o, hi[k].fx, hi[k].fy, hi[k].fz, hi[k].b);
        store_part(apart + o, hi[k].a);
      }
    }
  }
}

// The same pass with the j range of each slice split between the four waves
// of a workgroup, which share one 64-particle i tile: grid (ceil(N / 64), S).
// At small N the plain grid holds about one wave per SIMD or less (N = 2048:
// 8 x 32 workgroups) and the pass was latency-bound; here every slice runs as
// four waves.  Their sums are combined in LDS in wave order (deterministic)
// and written to the same partial layout.
template <bool PERIODIC, bool GUARDALL, bool COMP, bool TIES>
__global__ __launch_bounds__(kBlock) void k_force_split(const uint4* __restrict__ q,
                                                        const float4* __restrict__ pf,
                                                        float4* __restrict__ fpart,
                                                        float* __restrict__ apart, int N, int Npad,
                                                        int jchunk, float C1, float thr)
{
  __shared__ float red[kWaves - 1][5][64];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = blockIdx.x * 64;
  if (i0 >= N) return;  // a padding tile: the whole workgroup leaves
  const int s = blockIdx.y;
  const int js = s * jchunk;
  const int je = min(N, js + jchunk);
  const int jq = (jchunk + kWaves - 1) / kWaves;
  const int j0 = min(je, js + wv * jq);
  const int j1 = min(je, j0 + jq);
  PosT<PERIODIC> pi[1];
  int ii[1] = {i0 + lane};
  Acc hi[1], lo[1];
  pi[0] = load_pos<PERIODIC>(q, pf, min(ii[0], N - 1));
  acc_zero(hi[0]);
  acc_zero(lo[0]);
  if constexpr (GUARDALL) {
    j_range<PERIODIC, 2, 1, COMP, TIES>(q, pf, j0, j1, pi, ii, C1, thr, hi, lo);
  } else {
    const int d0 = max(j0, i0), d1 = min(j1, i0 + 64);
    if (d0 >= d1) {
      j_range<PERIODIC, 0, 1, COMP, TIES>(q, pf, j0, j1, pi, ii, C1, thr, hi, lo);
    } else {
      j_range<PERIODIC, 0, 1, COMP, TIES>(q, pf, j0, d0, pi, ii, C1, thr, hi, lo);
      j_range<PERIODIC, 1, 1, COMP, TIES>(q, pf, d0, d1, pi, ii, C1, thr, hi, lo);
      j_range<PERIODIC, 0, 1, COMP, TIES>(q, pf, d1, j1, pi, ii, C1, thr, hi, lo);
    }
  }
  float v[5];
  if constexpr (COMP) {
    v[0] = hi[0].fx + lo[0].fx;
    v[1] = hi[0].fy + lo[0].fy;
    v[2] = hi[0].fz + lo[0].fz;
    v[3] = hi[0].b + lo[0].b;
    v[4] = hi[0].a + lo[0].a;
  } else {
    v[0] = hi[0].fx;
    v[1] = hi[0].fy;
    v[2] = hi[0].fz;
    v[3] = hi[0].b;
    v[4] = hi[0].a;
  }
  if (wv > 0)
#pragma unroll
    for (int k = 0; k < 5; ++k) red[wv - 1][k][lane] = v[k];
  __syncthreads();
  if (wv != 0 || ii[0] >= N) return;
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int w = 0; w < kWaves - 1; ++w) v[k] += red[w][k][lane];
  const size_t o = (size_t)s * Npad + ii[0];
  store_part(fpart + o, v[0], v[1], v[2], v[3]);
  store_part(apart + o, v[4]);
}

__global__ __launch_bounds__(kBlock) void k_prep_q(const float4* __restrict__ pos,
                                                   uint4* __restrict__ q, int N, double qs)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const float4 p = pos[i];
  q[i] = make_uint4(fixq(p.x, qs), fixq(p.y, qs), fixq(p.z, qs), 0u);
}

__global__ __launch_bounds__(kBlock) void k_reduce_forces(const float4* __restrict__ fpart,
                                                          const float* __restrict__ apart,
                                                          float4* __restrict__ force,
                                                          double* __restrict__ blk, int ibase,
                                                          int iend, int S, int Npad, double FS)
{
  const int i = ibase + blockIdx.x * kBlock + threadIdx.x;
  double v[2] = {0., 0.};
  if (i < iend) {
    const Partial p = sum_slices(fpart, apart, i, S, Npad);
    const double e = 0.5 * (p.b - p.a);
    force[i] = make_float4((float)(FS * p.fx), (float)(FS * p.fy), (float)(FS * p.fz), (float)e);
    v[0] = e;
    v[1] = p.b;
  }
  const int fld[2] = {BLK_E, BLK_B};
  block_sum_store<2>(v, blk + (size_t)blockIdx.x * kBlkFields, fld);
}

__global__ __launch_bounds__(kBlock) void k_finalize_forces(const double* __restrict__ blk,
                                                            int nblk, double eps,
                                                            double* __restrict__ ev,
                                                            float* __restrict__ pressure_f)
{
  double v[2] = {0., 0.};
  for (int r = threadIdx.x; r < nblk; r += kBlock) {
    v[0] += blk[(size_t)r * kBlkFields + BLK_E];
    v[1] += blk[(size_t)r * kBlkFields + BLK_B];
  }
  block_allsum<2>(v);
  if (threadIdx.x == 0) {
    if (ev) {
      ev[0] = 2. * eps * v[0];   // E_pot = 4 eps sum_{i<j}
      ev[1] = 12. * eps * v[1];  // W = sum_{i<j} r.f
    }
    if (pressure_f) *pressure_f = (float)(4. * eps * v[1]);  // MDSystem.cu:108,136
  }
}

// ---------------------------------------------------------------------------
// integrator kernels: the reference's arithmetic (double expressions stored
// to float, MDSystem.cpp:447-505); FMA contraction is off so that given the
// same forces the device reproduces the reference's bits.

template <bool MAKEQ, int CANON>
__global__ __launch_bounds__(kBlock) void k_drift(StepArgs a)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.N) return;
  float4 p = a.pos[i];
  float4 v = a.vel[i];
  float4 tf;
  drift_one<CANON>(p, v, tf, a.force[i], a.dt);
  if constexpr (CANON == LJ_EVN) a.vel[i] = v;
  else a.tF[i] = tf;
  a.pos[i] = p;
  if constexpr (MAKEQ) a.q[i] = make_uint4(fixq(p.x, a.qscale), fixq(p.y, a.qscale), fixq(p.z, a.qscale), 0u);
}

// The kick kernels read the S partial entries of every particle, a sum bound
// by load latency: kSumLanes lanes share one particle's sum and a block holds
// kRowParticles particles, so that every CU takes part and many loads are in
// flight.  The lead lane of each particle does the integrator arithmetic; each
// block writes one block row.
// DRIFT: the kernel that ends a step also does the next step's drift for its
// particle (the arithmetic of k_drift on the values it just wrote), so that a
// run of steps needs no separate drift launch; MAKEQ: periodic fixed-point copy.
// FIN: block 0 computes the observables of the previous step of the run
// (`fin`: that step's arguments, whose block rows and force totals sit in the
// other of the two buffers) in place of a finaliser launch between the steps;
// the particle rows start at block 1.
constexpr int kSumLanes = 4;
constexpr int kSumBlock = kRowParticles * kSumLanes;
static_assert(kSumBlock == kBlock, "the row sums assume kBlock threads");

template <bool WITH_A, bool DRIFT, bool MAKEQ, bool FIN>
__global__ __launch_bounds__(kSumBlock) void k_kick_evn(StepArgs a, StepArgs fin)
{
#pragma clang fp contract(off)
  if constexpr (FIN) {
    if (blockIdx.x == 0) {
      step_observables<kSumBlock / 64>(fin, fin.nrow, true, true);
      return;
    }
  }
  const int row = (int)blockIdx.x - (FIN ? 1 : 0);
  const int r = threadIdx.x % kSumLanes;
  const int i = row * kRowParticles + threadIdx.x / kSumLanes;
  const bool lead = r == 0 && i < a.N;
  // the particle's own loads go out ahead of the slot sums (their latencies overlap)
  float4 v = {}, p = {};
  if (lead) {
    v = a.vel[i];
    if (DRIFT || a.bc != LJ_BC_NONE) p = a.pos[i];
  }
  const Partial pp = sum_slices_lanes<kSumLanes, WITH_A>(a.fpart, a.apart, i < a.N ? i : -1,
                                                         a.S, a.Npad, r);
  double v3[3] = {0., 0., 0.};
  if (lead) {
    const double e = 0.5 * (pp.b - pp.a);
    const float fx = (float)(a.FS * pp.fx), fy = (float)(a.FS * pp.fy), fz = (float)(a.FS * pp.fz);
    a.force[i] = make_float4(fx, fy, fz, (float)e);
    const double dt = a.dt;
    v.x = (float)(v.x + dt * fx / 2.);  // MDSystem.cpp:460-462
    v.y = (float)(v.y + dt * fy / 2.);
    v.z = (float)(v.z + dt * fz / 2.);
    if (a.bc != LJ_BC_NONE) {
      apply_bc_one(p, v, a.bc, a.L);
      if (!DRIFT && a.bc == LJ_BC_PERIODIC) a.pos[i] = p;
    }
    v3[0] = kin2(v);
    v3[1] = e;
    v3[2] = pp.b;
    if constexpr (DRIFT) {
      float4 tf;
      drift_one<LJ_EVN>(p, v, tf, make_float4(fx, fy, fz, 0.f), dt);
      a.pos[i] = p;
      if constexpr (MAKEQ)
        a.q[i] = make_uint4(fixq(p.x, a.qscale), fixq(p.y, a.qscale), fixq(p.z, a.qscale), 0u);
    }
    a.vel[i] = v;
  }
  const int fld[3] = {BLK_K2, BLK_E, BLK_B};
  block_sum_store<3>(v3, a.blk + (size_t)row * kBlkFields, fld);
}

template <bool WITH_A, bool FIN>
__global__ __launch_bounds__(kSumBlock) void k_tvn_a(StepArgs a, StepArgs fin)
{
#pragma clang fp contract(off)
  if constexpr (FIN) {
    if (blockIdx.x == 0) {
      step_observables<kSumBlock / 64>(fin, fin.nrow, true, true);
      return;
    }
  }
  const int row = (int)blockIdx.x - (FIN ? 1 : 0);
  const int r = threadIdx.x % kSumLanes;
  const int i = row * kRowParticles + threadIdx.x / kSumLanes;
  const bool lead = r == 0 && i < a.N;
  // the particle's own loads go out ahead of the slot sums (their latencies overlap)
  float4 tf = {}, v = {};
  if (lead) {
    tf = a.tF[i];
    v = a.vel[i];
  }
  const Partial pp = sum_slices_lanes<kSumLanes, WITH_A>(a.fpart, a.apart, i < a.N ? i : -1,
                                                         a.S, a.Npad, r);
  double v3[3] = {0., 0., 0.};
  if (lead) {
    const double e = 0.5 * (pp.b - pp.a);
    const float fx = (float)(a.FS * pp.fx), fy = (float)(a.FS * pp.fy), fz = (float)(a.FS * pp.fz);
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
    v3[2] = pp.b;
  }
  const int fld[3] = {BLK_TV2, BLK_E, BLK_B};
  block_sum_store<3>(v3, a.blk + (size_t)row * kBlkFields, fld);
}

// one particle per thread; each wave (kRowParticles particles) writes the K2
// field of its block row
template <bool DRIFT, bool MAKEQ>
__global__ __launch_bounds__(kBlock) void k_tvn_b(StepArgs a)
{
#pragma clang fp contract(off)
  static_assert(kRowParticles == 64, "one block row per wave");
  const int i = blockIdx.x * kBlock + threadIdx.x;
  // the particle's loads go out ahead of the row total (their latencies overlap)
  float4 v = {}, tf = {}, p = {}, f = {};
  if (i < a.N) {
    v = a.vel[i];
    tf = a.tF[i];
    if (DRIFT || a.bc != LJ_BC_NONE) p = a.pos[i];
    if (DRIFT) f = a.force[i];
  }
  const double tv2 = blk_total(a.blk, a.nrow, BLK_TV2);
  const double Tkin = tv2 * (1. / 3. / a.N);  // MDSystem.cpp:370
  const double chi = sqrt(a.T0 / Tkin);      // MDSystem.cpp:499
  double k2 = 0.;
  if (i < a.N) {
    const double dt = a.dt;
    v.x = (float)((2. * chi - 1.) * v.x + chi * dt * tf.x);  // MDSystem.cpp:503-505
    v.y = (float)((2. * chi - 1.) * v.y + chi * dt * tf.y);
    v.z = (float)((2. * chi - 1.) * v.z + chi * dt * tf.z);
    if (a.bc != LJ_BC_NONE) {
      apply_bc_one(p, v, a.bc, a.L);
      if (!DRIFT && a.bc == LJ_BC_PERIODIC) a.pos[i] = p;
    }
    a.vel[i] = v;
    k2 = kin2(v);
    if constexpr (DRIFT) {
      float4 tfn;
      drift_one<LJ_TVN>(p, v, tfn, f, dt);
      a.tF[i] = tfn;
      a.pos[i] = p;
      if constexpr (MAKEQ)
        a.q[i] = make_uint4(fixq(p.x, a.qscale), fixq(p.y, a.qscale), fixq(p.z, a.qscale), 0u);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) k2 += __shfl_xor(k2, o, 64);
  const int row = i / kRowParticles;
  if ((threadIdx.x & 63) == 0 && row < a.nrow) a.blk[(size_t)row * kBlkFields + BLK_K2] = k2;
}

__global__ __launch_bounds__(kBlock) void k_finalize_step(StepArgs a, int nblk)
{
  step_observables<kWaves>(a, nblk, true, true);
}

__global__ __launch_bounds__(kBlock) void k_params(StepArgs a, int nblk, int force_terms)
{
  step_observables<kWaves>(a, nblk, false, force_terms != 0);
}

__global__ void k_set_hist_base(lj_obs* obs, int64_t* base) { *base = obs->step; }

__global__ __launch_bounds__(kBlock) void k_kinetic_partials(const float4* __restrict__ vel, int N,
                                                             double* __restrict__ blk)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v1[1] = {0.};
  if (i < N) v1[0] = kin2(vel[i]);
  const int fld[1] = {BLK_K2};
  block_sum_store<1>(v1, blk + (size_t)blockIdx.x * kBlkFields, fld);
}

// RenormalizeVelocities (MDSystem.cpp:375-389): factor = sqrt(T0/Tkin) with
// Tkin from the K2 partials, or the given factor when factor > 0
__global__ __launch_bounds__(kBlock) void k_scale_velocities(float4* __restrict__ vel, int N,
                                                             const double* __restrict__ blk,
                                                             int nblk, double T0, double factor,
                                                             lj_obs* obs)
{
#pragma clang fp contract(off)
  double Tkin = 0.;
  if (!(factor > 0.)) {
    Tkin = blk_total(blk, nblk, BLK_K2) * (1. / 3. / N);
    factor = sqrt(T0 / Tkin);
  }
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < N) {
    float4 v = vel[i];
    v.x = (float)(v.x * factor);
    v.y = (float)(v.y * factor);
    v.z = (float)(v.z * factor);
    vel[i] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && obs) {
    if (Tkin > 0.) {
      obs->K *= T0 / Tkin;  // MDSystem.cpp:386-388
      obs->T = T0;
    } else {
      obs->K *= factor * factor;
      obs->T *= factor * factor;
    }
    obs->U = obs->K + obs->V;
  }
}

__global__ __launch_bounds__(kBlock) void k_apply_bc(float4* __restrict__ pos,
                                                     float4* __restrict__ vel, int N, int bc,
                                                     double L)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  float4 p = pos[i], v = vel[i];
  apply_bc_one(p, v, bc, L);
  pos[i] = p;
  vel[i] = v;
}

__global__ void k_store_vw(const double* ev, lj_obs* obs)
{
  obs->V = ev[0];
  obs->W = ev[1];
}

__global__ void k_reset_clock(lj_obs* obs)
{
  obs->t = 0.;
  obs->step = 0;
}

__global__ void k_reset_averaging(lj_obs* obs)
{
  obs->av_U_tot = obs->av_T_tot = obs->av_p_tot = 0.;  // MDSystem.cpp:701-705
  obs->av_iters = 0;
}

// ---------------------------------------------------------------------------
// RDF

// RDF pass.  Every binning rule is symmetric in i and j (the coordinate
// differences of the two orders are exact negatives, the image shift rounds
// half away from zero, r^2 squares them), so each unordered pair is binned
// once and counted twice: the ordered-pair counts of the reference, exactly.
// Work item (g, h), h >= g, over groups of kBlock particles (grid ng x ng;
// blocks with h < g exit at once); one i per thread.  SMALL (nbins <=
// kRdfWaveCopy): one LDS histogram per wave and group h staged in LDS; else one
// histogram per block and the j particles read from global memory, so that
// the LDS stays within 64 KiB at the largest nbins.  Bin divisions and the
// image rounding are table lookups (host-bisected thresholds), with the
// reference's float/double arithmetic reproduced.  Periodic boxes with SMALL
// histograms take k_rdf_fx, which decides almost every pair from the
// fixed-point image with a proven error bound and runs the reference
// arithmetic only near bin edges; the others take k_rdf.
constexpr int kRdfWaveCopy = 1024;

// The block's counts, one unit per unordered pair (doubled into the ordered
// counts when flushed), in one LDS histogram per wave, h[wave][bin].  Pairs of
// the last bin go to a register first: with the GPU rule every pair beyond the
// range is clamped there, which would put most lanes of a wave on one LDS
// address.  (A 32-column histogram per workgroup, h[bin][lane & 31], whose
// atomics never share a bank, measured slower on MI355X: 254 against 130 us
// at 500 bins, N = 16384, from 2 workgroups per CU at 64 KiB.)
__device__ __forceinline__ unsigned int* rdf_hist_init(unsigned int* hs, int nbins)
{
  for (int b = threadIdx.x; b < kWaves * nbins; b += kBlock) hs[b] = 0u;
  return hs + (int)(threadIdx.x >> 6) * nbins;
}

// adds the block's counts (with the waves' last-bin registers) to hist
__device__ __forceinline__ void rdf_flush(const unsigned int* hs, unsigned int* hw, int ncopy,
                                          int nbins, unsigned int last,
                                          unsigned long long* __restrict__ hist)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) last += __shfl_xor(last, o, 64);
  if ((threadIdx.x & 63) == 0 && last) atomicAdd(&hw[nbins - 1], last);
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += kBlock) {
    unsigned int c = 0u;
    for (int k = 0; k < ncopy; ++k) c += hs[k * nbins + b];
    if (c) atomicAdd(&hist[b], 2ull * c);
  }
}

template <int RULE, bool PERIODIC, bool SMALL>
__global__ __launch_bounds__(kBlock) void k_rdf(const float4* __restrict__ pos, int N, double L,
                                                int nbins, const float* __restrict__ tab,
                                                float est_scale, float t0, float t1, float t2,
                                                float t3, unsigned long long* __restrict__ hist)
{
  const int g = blockIdx.x, hgrp = blockIdx.y;
  if (hgrp < g) return;  // uniform, before any barrier
  extern __shared__ float4 sh_rdf4[];
  float4* pj_s = sh_rdf4;  // [kBlock] when SMALL
  unsigned int* hs = reinterpret_cast<unsigned int*>(sh_rdf4 + (SMALL ? kBlock : 0));
  float* T = reinterpret_cast<float*>(hs + (SMALL ? kWaves : 1) * nbins);
  unsigned int* hw = hs;
  if constexpr (SMALL) hw = rdf_hist_init(hs, nbins);
  else
    for (int b = threadIdx.x; b < nbins; b += kBlock) hs[b] = 0u;
  for (int b = threadIdx.x; b <= nbins; b += kBlock) T[b] = tab[b];
  const int j0 = hgrp * kBlock;
  const int nj = min(kBlock, N - j0);
  if constexpr (SMALL)
    if ((int)threadIdx.x < nj) pj_s[threadIdx.x] = pos[j0 + threadIdx.x];
  const float4* PJ = SMALL ? pj_s : pos + j0;
  __syncthreads();

  const int i = g * kBlock + threadIdx.x;
  const bool valid = i < N;
  const float4 pi = pos[min(i, N - 1)];
  const bool diag = g == hgrp;
  const float thr[4] = {t0, t1, t2, t3};
  const float Tmax = T[nbins], Tlast = T[nbins - 1];
  unsigned int last = 0u;
  if constexpr (!PERIODIC && SMALL) {
    // hard walls: the reference arithmetic of rdf_exact_bin without its
    // early returns, 8 pairs at a time (their LDS reads in flight together)
    const int jlo = valid ? (diag ? (int)threadIdx.x : -1) : nj;
    for (int j0b = 0; j0b < nj; j0b += 8) {
      constexpr int B = 8;
      float r2[B];
      int kk[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        r2[b] = rdf_wall_r2<RULE>(pi, PJ[min(j0b + b, nj - 1)]);
        const float e = RULE == LJ_RDF_RBINS ? __builtin_amdgcn_sqrtf(r2[b]) * est_scale
                                             : r2[b] * est_scale;
        kk[b] = rdf_bin(r2[b], T, nbins, (int)e);
      }
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int jj = j0b + b;
        // kept: r2 > 1e-6 (GPU rule) or below the last bin's end (the others)
        const bool keep = RULE == LJ_RDF_REF_GPU ? r2[b] > 1e-6f : r2[b] < Tmax;
        const bool count = (jj < nj) & (jj > jlo) & keep;
        const bool inlast = r2[b] >= Tlast;
        last += (count & inlast) ? 1u : 0u;
        if (count & !inlast) atomicAdd(&hw[kk[b]], 1u);
      }
    }
  } else {
#pragma unroll 4
    for (int jj = 0; jj < nj; ++jj) {
      if (!valid || (diag && jj <= (int)threadIdx.x)) continue;
      const int bin = rdf_exact_bin<RULE, PERIODIC>(pi, PJ[jj], L, (float)L, thr, T, nbins, Tmax,
                                                    Tlast, est_scale);
      if (bin == nbins - 1) ++last;
      else if (bin >= 0) atomicAdd(&hw[bin], 1u);
    }
  }
  rdf_flush(hs, hw, SMALL ? kWaves : 1, nbins, last, hist);  // !SMALL: the waves share one
}

// Periodic boxes, small histograms: every pair is first classified from the
// exact fixed-point minimum image (the force pass's, q = x 2^32 / L) under the
// error bound of RdfClass (lj_rdf_device.h), with A the span of the
// coordinates of the pair's two particle groups.  The pairs it leaves
// undecided -- about 2 delta / bin width of them, 1e-3 at 500 r-bins -- are
// marked in a bit mask per lane and walked through the reference arithmetic
// after each run of 32 j, so the exact path runs as often as the busiest lane
// needs it.
//   FAR (ranges that hold a small part of the box, as the reference rules'
// r2 < 25.6 at liquid densities): a first pass over 32 j marks the pairs
// that may lie in the range by one comparison of r2 with RdfClass's far
// bound; each lane then walks its marked pairs through the classification.
template <int RULE, bool FAR>
__global__ __launch_bounds__(kBlock) void k_rdf_fx(const float4* __restrict__ pos, int N, double L,
                                                   int nbins, const float* __restrict__ tab,
                                                   float est_scale, float t0, float t1, float t2,
                                                   float t3, double qscale,
                                                   unsigned long long* __restrict__ hist)
{
  constexpr bool TAB = RULE != LJ_RDF_RBINS;  // the classification reads (T[k], T[k + 1])
  const int g = blockIdx.x, hgrp = blockIdx.y;
  if (hgrp < g) return;  // uniform, before any barrier
  // the j group apart from the histograms, so that the loads of a batch of j
  // may pass the atomics of the previous one
  __shared__ float4 pj_s[kBlock];
  __shared__ uint4 qj_s[kBlock];
  __shared__ float bw[2][kWaves];
  // dynamic LDS: T2 [nbins] when TAB: (T[k], T[k + 1]) S^2; T [nbins + 1]; the
  // per-wave histograms
  extern __shared__ float2 sh_rdf2[];
  float2* T2 = sh_rdf2;
  float* T = reinterpret_cast<float*>(T2 + (TAB ? nbins : 0));
  unsigned int* hs = reinterpret_cast<unsigned int*>(T + nbins + 1);
  const int wv = (int)(threadIdx.x >> 6);
  unsigned int* hw = rdf_hist_init(hs, nbins);
  const double S2 = qscale * qscale;
  for (int b = threadIdx.x; b <= nbins; b += kBlock) {
    const float t = tab[b];
    T[b] = t;
    if constexpr (TAB) {
      const float tq = (float)((double)t * S2);
      if (b < nbins) T2[b].x = tq;
      if (b > 0) T2[b - 1].y = tq;
    }
  }
  const int j0 = hgrp * kBlock;
  const int nj = min(kBlock, N - j0);
  float cmax = -3.0e38f, cmin = 3.0e38f;
  if ((int)threadIdx.x < nj) {
    const float4 p = pos[j0 + threadIdx.x];
    pj_s[threadIdx.x] = p;
    qj_s[threadIdx.x] = make_uint4(fixq(p.x, qscale), fixq(p.y, qscale), fixq(p.z, qscale), 0u);
    cmax = fmaxf(p.x, fmaxf(p.y, p.z));
    cmin = fminf(p.x, fminf(p.y, p.z));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    cmax = fmaxf(cmax, __shfl_xor(cmax, o, 64));
    cmin = fminf(cmin, __shfl_xor(cmin, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    bw[0][wv] = cmax;
    bw[1][wv] = cmin;
  }
  __syncthreads();

  const int i = g * kBlock + threadIdx.x;
  const float4 pi = pos[min(i, N - 1)];
  const uint4 qi = make_uint4(fixq(pi.x, qscale), fixq(pi.y, qscale), fixq(pi.z, qscale), 0u);
  // this lane's pairs: j > jlo (the diagonal group counts each pair once)
  const int jlo = i < N ? (g == hgrp ? (int)threadIdx.x : -1) : kBlock;
  const float thr[4] = {t0, t1, t2, t3};
  const float Tmax = T[nbins], Tlast = T[nbins - 1];
  // the bound's span A: the coordinates of this lane's particle and the j group
  float hmax = fmaxf(pi.x, fmaxf(pi.y, pi.z)), hmin = fminf(pi.x, fminf(pi.y, pi.z));
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    hmax = fmaxf(hmax, bw[0][w]);
    hmin = fminf(hmin, bw[1][w]);
  }
  RdfClass<RULE> cl;
  cl.init((double)hmax - (double)hmin, L, qscale, est_scale, nbins, Tlast, Tmax);
  unsigned int last = 0u;

  // the reference arithmetic for pair (i, jj), counted
  auto exact = [&](int jj) {
    const int bin = rdf_exact_bin<RULE, true>(pi, pj_s[jj], L, (float)L, thr, T, nbins, Tmax, Tlast,
                                              est_scale);
    if (bin == nbins - 1) ++last;
    else if (bin >= 0) atomicAdd(&hw[bin], 1u);
  };
  // one pair, decided or walked to the exact path at once (the walks)
  auto one = [&](int jj) {
    const float r2 = rdf_r2q(qi, qj_s[jj]);
    int k;
    bool dec, gone;
    if constexpr (TAB) {
      float lo, hi;
      cl.measure(r2, k, lo, hi);
      cl.decide(k, lo, hi, T2[k], dec, gone);
    } else {
      cl.rbins(r2, k, dec, gone);
    }
    if (dec) atomicAdd(&hw[k], 1u);
    else if (gone) {
      if constexpr (RULE == LJ_RDF_REF_GPU) ++last;
    } else {
      exact(jj);
    }
  };

  for (int c0 = 0; c0 < nj; c0 += 32) {  // kBlock is a multiple of 32
    unsigned int mask = 0u;
    if constexpr (FAR) {
      const float X = cl.far_bound();
#pragma unroll 8
      for (int b = 0; b < 32; ++b) {
        const int jj = c0 + b;
        const float r2 = rdf_r2q(qi, qj_s[jj]);
        const bool ok = (jj > jlo) & (jj < nj);
        const bool far = r2 >= X;
        if constexpr (RULE == LJ_RDF_REF_GPU) last += (ok & far) ? 1u : 0u;
        mask |= (ok & !far) ? (1u << b) : 0u;
      }
      while (mask) {
        const int b = __builtin_ctz(mask);
        mask &= mask - 1u;
        one(c0 + b);
      }
    } else {
      constexpr int B = 8;
#pragma unroll 1
      for (int b0 = 0; b0 < 32; b0 += B) {
        int kk[B];
        bool dec[B], gone[B];
        if constexpr (TAB) {
          float lo[B], hi[B];
          float2 t[B];
#pragma unroll
          for (int b = 0; b < B; ++b) {
            cl.measure(rdf_r2q(qi, qj_s[c0 + b0 + b]), kk[b], lo[b], hi[b]);
            t[b] = T2[kk[b]];
          }
#pragma unroll
          for (int b = 0; b < B; ++b) cl.decide(kk[b], lo[b], hi[b], t[b], dec[b], gone[b]);
        } else {
#pragma unroll
          for (int b = 0; b < B; ++b) cl.rbins(rdf_r2q(qi, qj_s[c0 + b0 + b]), kk[b], dec[b], gone[b]);
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const int jj = c0 + b0 + b;
          const bool ok = (jj > jlo) & (jj < nj);
          if constexpr (RULE == LJ_RDF_REF_GPU) last += (ok & gone[b]) ? 1u : 0u;
          if (ok & dec[b]) atomicAdd(&hw[kk[b]], 1u);
          mask |= (ok & !dec[b] & !gone[b]) ? (1u << (b0 + b)) : 0u;
        }
      }
      while (mask) {
        const int b = __builtin_ctz(mask);
        mask &= mask - 1u;
        exact(c0 + b);
      }
    }
  }
  rdf_flush(hs, hw, kWaves, nbins, last, hist);
}

// speed histogram of MDSystem::initvelo / updatevelo (MDSystem.cpp:651-694):
// bin = (int)(sqrt((double)fl32(vx vx + vy vy + vz vz)) / step), kept below
// nbins; LDS counters per workgroup, one 64-bit atomic per bin
__global__ __launch_bounds__(kBlock) void k_velocity_hist(const float4* __restrict__ vel, int N,
                                                          double step, int nbins,
                                                          unsigned long long* __restrict__ hist)
{
#pragma clang fp contract(off)
  extern __shared__ unsigned int sh_vh[];
  for (int b = threadIdx.x; b < nbins; b += kBlock) sh_vh[b] = 0u;
  __syncthreads();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
    const float4 v = vel[i];
    const float s = v.x * v.x + v.y * v.y + v.z * v.z;
    const double t = __builtin_sqrt((double)s) / step;
    if (t < (double)nbins) atomicAdd(&sh_vh[(int)t], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += kBlock)
    if (sh_vh[b]) atomicAdd(&hist[b], (unsigned long long)sh_vh[b]);
}

// Shear stress of the reference CPU path.  CalculateForces sums -r_x f_y over
// the ordered pairs (MDSystem.cpp:299) and scales the sum by 4/2 (:309);
// CalculateParameters adds -v_x v_y per particle (:335) and divides by the
// volume N / rho (:353).  The pair sum takes the force pass's FP32 pair
// arithmetic (the fixed-point minimum image in periodic boxes) over the
// sliced grid (i blocks x j slices), in FP32 blocks of kJB j added in FP64:
// part[workgroup] = sum of -dx dy s over its pairs (internal units).  j is the
// same for every lane (scalar loads).
template <bool PERIODIC>
__global__ __launch_bounds__(kBlock) void k_shear_pairs(const uint4* __restrict__ q,
                                                        const float4* __restrict__ pf, int N,
                                                        int jchunk, float C1,
                                                        double* __restrict__ part)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int j0 = blockIdx.y * jchunk, j1 = min(N, j0 + jchunk);
  const PosT<PERIODIC> pi = load_pos<PERIODIC>(q, pf, min(i, N - 1));
  double v[1] = {0.};
  for (int b0 = j0; b0 < j1; b0 += kJB) {
    const int b1 = min(j1, b0 + kJB);
    float t = 0.f;
    for (int j = b0; j < b1; ++j) {
      const PosT<PERIODIC> pj = load_pos<PERIODIC>(q, pf, j);
      float dx, dy, dz;
      if constexpr (PERIODIC) {
        dx = mi_fixed<true>(pi.x, pj.x);
        dy = mi_fixed<true>(pi.y, pj.y);
        dz = mi_fixed<true>(pi.z, pj.z);
      } else {
        dx = pi.x - pj.x;
        dy = pi.y - pj.y;
        dz = pi.z - pj.z;
      }
      const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
      float ir2 = __builtin_amdgcn_rcpf(r2) * C1;
      ir2 = j == i ? 0.f : ir2;  // the reference's j != i
      const float e6 = ir2 * ir2 * ir2;
      const float s = e6 * __builtin_fmaf(e6, 2.f, -1.f) * ir2;
      t = __builtin_fmaf(-dx * dy, s, t);
    }
    v[0] += (double)t;
  }
  if (i >= N) v[0] = 0.;
  const int fld[1] = {0};
  block_sum_store<1>(v, part + (size_t)blockIdx.y * gridDim.x + blockIdx.x, fld);
}

// out = (scale * sum of the pair parts + sum_i fl32(-v_x v_y)) / volume, both
// sums in a fixed order; scale = 1/2 (the ordered pairs' -r_x f_y, times 4/2,
// with the reference's force at a quarter of f) times the force pass's
// internal-unit factor
__global__ __launch_bounds__(kBlock) void k_shear_finalize(const double* __restrict__ part,
                                                           int nparts,
                                                           const float4* __restrict__ vel, int N,
                                                           double scale, double volume,
                                                           double* __restrict__ out)
{
#pragma clang fp contract(off)
  double v[2] = {0., 0.};
  for (int r = threadIdx.x; r < nparts; r += kBlock) v[0] += part[r];
  for (int i = threadIdx.x; i < N; i += kBlock) {
    const float4 u = vel[i];
    const float ps = -u.x * u.y;  // MDSystem.cpp:335
    v[1] += (double)ps;
  }
  block_allsum<2>(v);
  if (threadIdx.x == 0) *out = (scale * v[0] + v[1]) / volume;
}

__global__ void k_fill_u64(unsigned long long* p, int n)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

inline int nblocks(int n) { return (n + kBlock - 1) / kBlock; }

// the j-split pass serves one-particle-per-lane plans whose plain grid holds
// under four waves per SIMD (1024 SIMDs); LJ_FORCE_SPLIT=0/1 forces it off or on
bool force_split_waves(const ForcePlan& pl)
{
  static const char* e = std::getenv("LJ_FORCE_SPLIT");
  if (pl.ipl != 1 || pl.n3) return false;
  if (e && e[0] == '0') return false;
  if (e && e[0] == '1') return true;
  return (long)pl.gx * kWaves * pl.S < 4L * 1024;
}

template <bool PERIODIC, bool GUARDALL, int IPL, bool COMP, bool TIES>
hipError_t launch_force_t(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                          const uint4* q, const float4* posf, float4* fpart, float* apart)
{
  if constexpr (IPL == 1) {
    if (!pl.sharded() && force_split_waves(pl)) {
      // tiles padded to a multiple of the 8 XCDs: tile t is written on XCD
      // t mod 8 by every slice's workgroup, as the 64-particle row t of the
      // integrator kernels that read its partials is scheduled (round robin)
      dim3 grid((unsigned)((((pl.N + 63) / 64) + 7) / 8 * 8), pl.S);
      hipExtLaunchKernelGGL((k_force_split<PERIODIC, GUARDALL, COMP, TIES>), grid, dim3(kBlock), 0,
                            st, tl_force_stamp.start, tl_force_stamp.stop, 0, q, posf, fpart, apart,
                            pl.N, pl.Npad, pl.jchunk, ps.C1, ps.thr);
      return hipGetLastError();
    }
  }
  dim3 grid(pl.gx, pl.S);
  hipExtLaunchKernelGGL((k_force<PERIODIC, GUARDALL, IPL, COMP, TIES>), grid, dim3(kBlock), 0, st,
                        tl_force_stamp.start, tl_force_stamp.stop, 0, q, posf, fpart, apart, pl.N,
                        pl.Npad, pl.jchunk, ps.C1, ps.thr, pl.i0, pl.iend());
  return hipGetLastError();
}

template <bool PERIODIC, bool GUARDALL, bool TIES>
hipError_t launch_force_c(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                          const uint4* q, const float4* posf, float4* fpart, float* apart,
                          bool comp)
{
  if (pl.ipl == 2)
    return comp ? launch_force_t<PERIODIC, GUARDALL, 2, true, TIES>(st, pl, ps, q, posf, fpart, apart)
                : launch_force_t<PERIODIC, GUARDALL, 2, false, TIES>(st, pl, ps, q, posf, fpart, apart);
  return comp ? launch_force_t<PERIODIC, GUARDALL, 1, true, TIES>(st, pl, ps, q, posf, fpart, apart)
              : launch_force_t<PERIODIC, GUARDALL, 1, false, TIES>(st, pl, ps, q, posf, fpart, apart);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers

ForcePlan make_force_plan(int N, int jslices_override, int ipl)
{
  ForcePlan pl;
  pl.N = N;
  pl.ipl = (ipl == 2) ? 2 : 1;
  const int waves_i = (N + 64 * pl.ipl - 1) / (64 * pl.ipl);
  pl.gx = (waves_i + kWaves - 1) / kWaves;
  int S;
  if (jslices_override > 0) {
    S = jslices_override;
  } else {
    // about 8 resident waves per SIMD on 256 CUs x 4 SIMDs
    const int target = 8192;
    S = (target + pl.gx * kWaves - 1) / (pl.gx * kWaves);
    const int smax = (N + 63) / 64;  // at least 64 j per slice
    S = max(1, min(S, smax));
  }
  S = max(1, min(S, N));
  int jc = (N + S - 1) / S;
  jc = (jc + 7) & ~7;
  pl.jchunk = jc;
  pl.S = (N + jc - 1) / jc;
  pl.Npad = (N + 63) & ~63;
  return pl;
}

ForcePlan make_shard_plan(int N, int shard, int nshards, int jslices)
{
  // whole 64-particle tiles (the kernel's wave tiles), so that every wave of
  // a share splits its slices' j ranges where the full pass does (the self
  // pair's segment) and takes the same FP32 blocks: bit-identical rows
  const long T = (N + 63) / 64;
  const int i0 = (int)(64 * (shard * T / nshards));
  const int i1 = (int)std::min<long>(N, 64 * ((shard + 1) * T / nshards));
  const int gx = (i1 - i0 + kBlock - 1) / kBlock;
  int S = jslices;
  if (S <= 0) {
    // about 8 resident waves per SIMD on 256 CUs x 4 SIMDs, as for a full pass
    S = (8192 + gx * kWaves - 1) / (gx * kWaves);
    S = max(1, min(S, (N + 63) / 64));
  }
  ForcePlan pl = make_force_plan(N, S, 1);  // the full pass's slicing for that count
  pl.i0 = i0;
  pl.i1 = i1;
  pl.gx = gx;
  return pl;
}

PairScalars make_pair_scalars(double L, int periodic, double eps, double sigma, bool guard_all)
{
  PairScalars ps;
  ps.periodic = periodic;
  ps.L = L;
  ps.eps = eps;
  ps.sigma = sigma;
  const double s2 = sigma * sigma;
  if (periodic) {
    // one fixed-point unit is u = L / 2^32: the box spans exactly 2^32 units
    const double u = L / 4294967296.0;
    ps.qscale = 4294967296.0 / L;
    ps.C1 = (float)(s2 / (u * u));
    ps.FS = 24. * eps * u / s2;
    ps.thr = guard_all ? (float)(1e-6 / (u * u)) : 0.f;
  } else {
    ps.qscale = 0.;
    ps.C1 = (float)s2;
    ps.FS = 24. * eps / s2;
    ps.thr = guard_all ? 1e-6f : 0.f;
  }
  return ps;
}

hipError_t launch_prep_q(hipStream_t st, const float4* pos, uint4* q, int N, double qscale)
{
  hipLaunchKernelGGL(k_prep_q, dim3(nblocks(N)), dim3(kBlock), 0, st, pos, q, N, qscale);
  return hipGetLastError();
}

hipError_t launch_force(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                        const uint4* q, const float4* posf, float4* fpart, float* apart,
                        bool compensated, bool ties, double* step_ew)
{
  const bool ga = ps.thr > 0.f;
  // the pair-once pass of the matching mode (plans are made per context)
  if (pl.n3 && compensated == (pl.comp != 0) && !ga)
    return launch_force_n3(st, pl, ps, q, posf, fpart, apart, ties, step_ew);
  if (ps.periodic) {
    if (ties)
      return ga ? launch_force_c<true, true, true>(st, pl, ps, q, posf, fpart, apart, compensated)
                : launch_force_c<true, false, true>(st, pl, ps, q, posf, fpart, apart, compensated);
    return ga ? launch_force_c<true, true, false>(st, pl, ps, q, posf, fpart, apart, compensated)
              : launch_force_c<true, false, false>(st, pl, ps, q, posf, fpart, apart, compensated);
  }
  return ga ? launch_force_c<false, true, false>(st, pl, ps, q, posf, fpart, apart, compensated)
            : launch_force_c<false, false, false>(st, pl, ps, q, posf, fpart, apart, compensated);
}

hipError_t launch_reduce_forces(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                                const float4* fpart, const float* apart, float4* force,
                                double* blk)
{
  hipLaunchKernelGGL(k_reduce_forces, dim3(nblocks(pl.iend() - pl.i0)), dim3(kBlock), 0, st, fpart,
                     apart, force, blk, pl.i0, pl.iend(), pl.S, pl.Npad, ps.FS);
  return hipGetLastError();
}

hipError_t launch_finalize_forces(hipStream_t st, const double* blk, int nblk, double eps,
                                  double* ev, float* pressure_f)
{
  hipLaunchKernelGGL(k_finalize_forces, dim3(1), dim3(kBlock), 0, st, blk, nblk, eps, ev,
                     pressure_f);
  return hipGetLastError();
}

hipError_t launch_drift(hipStream_t st, const StepArgs& a)
{
  const bool mq = a.bc == LJ_BC_PERIODIC;
  dim3 g(nblocks(a.N)), b(kBlock);
  if (a.canonical == LJ_EVN) {
    if (mq) hipLaunchKernelGGL((k_drift<true, LJ_EVN>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_drift<false, LJ_EVN>), g, b, 0, st, a);
  } else {
    if (mq) hipLaunchKernelGGL((k_drift<true, LJ_TVN>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_drift<false, LJ_TVN>), g, b, 0, st, a);
  }
  return hipGetLastError();
}

template <bool WITH_A, bool FIN>
void kick_evn_w(hipStream_t st, const StepArgs& a, bool drift_next, const StepArgs& fin)
{
  const dim3 g(a.nrow + (FIN ? 1 : 0)), b(kSumBlock);
  if (!drift_next)
    hipLaunchKernelGGL((k_kick_evn<WITH_A, false, false, FIN>), g, b, 0, st, a, fin);
  else if (a.bc == LJ_BC_PERIODIC)
    hipLaunchKernelGGL((k_kick_evn<WITH_A, true, true, FIN>), g, b, 0, st, a, fin);
  else
    hipLaunchKernelGGL((k_kick_evn<WITH_A, true, false, FIN>), g, b, 0, st, a, fin);
}

hipError_t launch_kick_evn(hipStream_t st, const StepArgs& a, bool drift_next, const StepArgs* fin)
{
  if (a.nitems) {  // force-only partials of the pair-once time-step pass
    if (fin)
      kick_evn_w<false, true>(st, a, drift_next, *fin);
    else
      kick_evn_w<false, false>(st, a, drift_next, a);
  } else {
    if (fin) return hipErrorInvalidValue;  // the observables block rides on the pair-once step only
    kick_evn_w<true, false>(st, a, drift_next, a);
  }
  return hipGetLastError();
}

hipError_t launch_tvn_a(hipStream_t st, const StepArgs& a, const StepArgs* fin)
{
  if (a.nitems) {
    if (fin)
      hipLaunchKernelGGL((k_tvn_a<false, true>), dim3(a.nrow + 1), dim3(kSumBlock), 0, st, a, *fin);
    else
      hipLaunchKernelGGL((k_tvn_a<false, false>), dim3(a.nrow), dim3(kSumBlock), 0, st, a, a);
  } else {
    if (fin) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_tvn_a<true, false>), dim3(a.nrow), dim3(kSumBlock), 0, st, a, a);
  }
  return hipGetLastError();
}

hipError_t launch_tvn_b(hipStream_t st, const StepArgs& a, bool drift_next)
{
  const dim3 g(nblocks(a.N)), b(kBlock);
  if (!drift_next) hipLaunchKernelGGL((k_tvn_b<false, false>), g, b, 0, st, a);
  else if (a.bc == LJ_BC_PERIODIC) hipLaunchKernelGGL((k_tvn_b<true, true>), g, b, 0, st, a);
  else hipLaunchKernelGGL((k_tvn_b<true, false>), g, b, 0, st, a);
  return hipGetLastError();
}

hipError_t launch_finalize_step(hipStream_t st, const StepArgs& a)
{
  hipLaunchKernelGGL(k_finalize_step, dim3(1), dim3(kBlock), 0, st, a, a.nrow);
  return hipGetLastError();
}

hipError_t launch_params(hipStream_t st, const StepArgs& a, int include_force_terms)
{
  hipLaunchKernelGGL(k_params, dim3(1), dim3(kBlock), 0, st, a, nblocks(a.N),
                     include_force_terms);
  return hipGetLastError();
}

hipError_t launch_set_hist_base(hipStream_t st, lj_obs* obs, int64_t* base)
{
  hipLaunchKernelGGL(k_set_hist_base, dim3(1), dim3(1), 0, st, obs, base);
  return hipGetLastError();
}

hipError_t launch_kinetic_partials(hipStream_t st, const float4* vel, int N, double* blk)
{
  hipLaunchKernelGGL(k_kinetic_partials, dim3(nblocks(N)), dim3(kBlock), 0, st, vel, N, blk);
  return hipGetLastError();
}

hipError_t launch_scale_velocities(hipStream_t st, float4* vel, int N, const double* blk,
                                   int nblk, double T0, double factor, lj_obs* obs)
{
  hipLaunchKernelGGL(k_scale_velocities, dim3(nblocks(N)), dim3(kBlock), 0, st, vel, N, blk, nblk,
                     T0, factor, obs);
  return hipGetLastError();
}

hipError_t launch_apply_bc(hipStream_t st, float4* pos, float4* vel, int N, int bc, double L)
{
  hipLaunchKernelGGL(k_apply_bc, dim3(nblocks(N)), dim3(kBlock), 0, st, pos, vel, N, bc, L);
  return hipGetLastError();
}

hipError_t launch_reset_averaging(hipStream_t st, lj_obs* obs)
{
  hipLaunchKernelGGL(k_reset_averaging, dim3(1), dim3(1), 0, st, obs);
  return hipGetLastError();
}

hipError_t launch_store_vw(hipStream_t st, const double* ev, lj_obs* obs)
{
  hipLaunchKernelGGL(k_store_vw, dim3(1), dim3(1), 0, st, ev, obs);
  return hipGetLastError();
}

hipError_t launch_reset_clock(hipStream_t st, lj_obs* obs)
{
  hipLaunchKernelGGL(k_reset_clock, dim3(1), dim3(1), 0, st, obs);
  return hipGetLastError();
}

hipError_t launch_rdf(hipStream_t st, const RdfArgs& a)
{
  const int ng = nblocks(a.N);
  const bool small = a.nbins <= kRdfWaveCopy;
  const dim3 grid(ng, ng), blk(kBlock);
  const size_t tab = sizeof(float) * (a.nbins + 1);
  if (small && rdf_uses_cells(a)) return launch_rdf_cells(st, a);
  if (small && a.periodic) {
    const size_t lds = (a.rule != LJ_RDF_RBINS ? sizeof(float2) * a.nbins : 0) + tab +
                       sizeof(unsigned) * kWaves * a.nbins;
    // the far-pair pass where the counted range holds under 30% of the box
    // (a sphere of radius sqrt(end) against L^3)
    const double re = std::sqrt(std::fmax(a.end, 0.));
    const bool far = re < 0.5 * a.L && 4.18879020478639098 * re * re * re < 0.3 * a.L * a.L * a.L;
#define LJ_RDF_FX_ARGS a.pos, a.N, a.L, a.nbins, a.tab, a.est_scale, a.thr[0], a.thr[1], a.thr[2], \
                       a.thr[3], 4294967296. / a.L, a.hist
#define LJ_RDF_FX(R)                                                                       \
  do {                                                                                     \
    if (far) hipLaunchKernelGGL((k_rdf_fx<R, true>), grid, blk, lds, st, LJ_RDF_FX_ARGS); \
    else hipLaunchKernelGGL((k_rdf_fx<R, false>), grid, blk, lds, st, LJ_RDF_FX_ARGS);    \
  } while (0)
    if (a.rule == LJ_RDF_REF_CPU) LJ_RDF_FX(LJ_RDF_REF_CPU);
    else if (a.rule == LJ_RDF_REF_GPU) LJ_RDF_FX(LJ_RDF_REF_GPU);
    else LJ_RDF_FX(LJ_RDF_RBINS);
#undef LJ_RDF_FX
#undef LJ_RDF_FX_ARGS
    return hipGetLastError();
  }
  const size_t lds =
      (small ? sizeof(float4) * kBlock + sizeof(unsigned) * kWaves * a.nbins : sizeof(unsigned) * a.nbins) + tab;
#define LJ_RDF_ARGS a.pos, a.N, a.L, a.nbins, a.tab, a.est_scale, a.thr[0], a.thr[1], a.thr[2], \
                    a.thr[3], a.hist
#define LJ_RDF_LAUNCH(R)                                                                             \
  do {                                                                                               \
    if (small) hipLaunchKernelGGL((k_rdf<R, false, true>), grid, blk, lds, st, LJ_RDF_ARGS);        \
    else if (a.periodic) hipLaunchKernelGGL((k_rdf<R, true, false>), grid, blk, lds, st, LJ_RDF_ARGS); \
    else hipLaunchKernelGGL((k_rdf<R, false, false>), grid, blk, lds, st, LJ_RDF_ARGS);             \
  } while (0)
  if (a.rule == LJ_RDF_REF_CPU) LJ_RDF_LAUNCH(LJ_RDF_REF_CPU);
  else if (a.rule == LJ_RDF_REF_GPU) LJ_RDF_LAUNCH(LJ_RDF_REF_GPU);
  else LJ_RDF_LAUNCH(LJ_RDF_RBINS);
#undef LJ_RDF_LAUNCH
#undef LJ_RDF_ARGS
  return hipGetLastError();
}

hipError_t launch_velocity_hist(hipStream_t st, const float4* vel, int N, double step, int nbins,
                                unsigned long long* hist)
{
  const int g = max(1, min(nblocks(N), 512));
  hipLaunchKernelGGL(k_velocity_hist, dim3(g), dim3(kBlock), sizeof(unsigned) * nbins, st, vel, N,
                     step, nbins, hist);
  return hipGetLastError();
}

hipError_t launch_shear(hipStream_t st, const ForcePlan& pl, const PairScalars& ps,
                        const uint4* q, const float4* pos, const float4* vel, double volume,
                        double* part, double* out)
{
  const dim3 grid(pl.gx, pl.S);
  if (ps.periodic)
    hipLaunchKernelGGL(k_shear_pairs<true>, grid, dim3(kBlock), 0, st, q, pos, pl.N, pl.jchunk,
                       ps.C1, part);
  else
    hipLaunchKernelGGL(k_shear_pairs<false>, grid, dim3(kBlock), 0, st, q, pos, pl.N, pl.jchunk,
                       ps.C1, part);
  // -r_x f_y = -(u dx) (FS s dy): u = L / 2^32 in periodic boxes, 1 otherwise
  const double u = ps.periodic ? ps.L / 4294967296.0 : 1.0;
  hipLaunchKernelGGL(k_shear_finalize, dim3(1), dim3(kBlock), 0, st, part, pl.gx * pl.S, vel, pl.N,
                     0.5 * u * ps.FS, volume, out);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t st, unsigned long long* p, int n)
{
  hipLaunchKernelGGL(k_fill_u64, dim3(nblocks(n)), dim3(kBlock), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace lj
