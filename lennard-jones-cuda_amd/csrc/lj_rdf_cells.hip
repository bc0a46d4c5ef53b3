// lj_rdf_cells.hip -- the RDF pass over a cell list, for periodic boxes whose
// counted range is short against the box (the reference rules' r^2 < 25.6 at
// liquid densities: N = 16384, rho* = 0.85 has 10 cells per axis at reach 2).
//
// The particles are sorted into nc^3 cells of side s >= sqrt(X) (reach 1)
// or s >= sqrt(X) / 2 (reach 2), X the classification's far bound
// (lj_rdf_device.h).  Every pair whose fixed-point r2 is below X lies in the
// same or adjacent cells (within two cells on each axis at reach 2).  Every
// other pair is, by the error bound, beyond the counted range in the
// reference's arithmetic too: dropped (CPU rule, r-bins) or in the clamped
// last bin (GPU rule).
//   Reach 1 (k_rdf_cells): a work item is (cell c, forward offset o), the
//   cell's own pairs once (o = 0) or its pairs with one of the 13 forward
//   neighbours (half a shell), so every unordered pair of adjacent cells is
//   met once (nc >= 3 makes the 27 neighbours distinct).
//   Reach 2 (k_rdf_cells2): one workgroup per cell against itself and its 62
//   forward neighbours (nc >= 5).
// The pairs are classified as in k_rdf_fx, with the reference arithmetic for
// the pairs near a bin edge.  The GPU rule's clamp bin then holds the
// N (N - 1) / 2 pairs less those counted elsewhere (or excluded at
// r^2 <= 1e-6).
//
// The bound holds for coordinate spans up to A = 3 L (all coordinates in
// [-L, 2 L), as the periodic wrap keeps them); the assignment kernel flags
// any other coordinate, the pass then counts nothing, and the flag is left
// for the host (lj_rdf falls back to the all-pairs pass).
//
// Kernels: the sort -- up to 4096 cells, k_cells_assign_lds (cell and rank
// in block-private LDS counts, the block's offset in each cell) and
// k_cells_place (the cells' starts, the sorted fixed-point positions with
// their index); else assign (global counters), scan (one block) and scatter
// -- then the pairs and the reduction.  Counts are integers added by atomics
// (order-independent), so the histogram is exactly the reference's whatever
// the order within a cell.  Workgroup histograms go to one of kRdfReps
// replicas, which the reduction sums, doubles into ordered-pair counts, adds
// to hist and clears for the next pass (no memset node in a captured step).
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "lj_internal.h"
#include "lj_rdf_device.h"
#include "lj_util.h"

namespace lj {

namespace {

using namespace dev;
using namespace rdfdev;

constexpr int kCellBlock = 256;
constexpr int kJStage = 512;  // neighbour-cell particles staged in LDS at a time

__device__ __forceinline__ int cell_of(const uint4& q, int nc)
{
  const int cx = (int)(((unsigned long long)q.x * (unsigned)nc) >> 32);
  const int cy = (int)(((unsigned long long)q.y * (unsigned)nc) >> 32);
  const int cz = (int)(((unsigned long long)q.z * (unsigned)nc) >> 32);
  return cx + nc * (cy + nc * cz);
}

// A particle's cell and its sorted record.  Periodic boxes: the cell of the
// fixed-point coordinates, the record (q.x, q.y, q.z, index), and `bad` when
// a coordinate lies outside [-L, 2 L) (the bound's span).  Hard walls
// (WALL): floor(x nc / L) per axis clamped to the grid -- particles a little
// outside [0, L] go to the edge cells, and the map stays monotone, so a pair
// closer than a cell's side is at most one cell apart per axis -- and the
// record holds the float coordinates' bits.
template <bool WALL>
__device__ __forceinline__ int cell_record(const float4& p, int i, double L, double qs, float inv_side,
                                           int nc, uint4& rec, bool& bad)
{
  if constexpr (WALL) {
    rec = make_uint4(__float_as_uint(p.x), __float_as_uint(p.y), __float_as_uint(p.z), (unsigned)i);
    bad = false;
    const int cx = min(max((int)floorf(p.x * inv_side), 0), nc - 1);
    const int cy = min(max((int)floorf(p.y * inv_side), 0), nc - 1);
    const int cz = min(max((int)floorf(p.z * inv_side), 0), nc - 1);
    return cx + nc * (cy + nc * cz);
  } else {
    rec = make_uint4(fixq(p.x, qs), fixq(p.y, qs), fixq(p.z, qs), (unsigned)i);
    const double lo = -L, hi = 2. * L;
    bad = !((double)p.x >= lo && (double)p.x < hi && (double)p.y >= lo && (double)p.y < hi &&
            (double)p.z >= lo && (double)p.z < hi);
    return cell_of(rec, nc);
  }
}

template <bool WALL>
__global__ __launch_bounds__(kCellBlock) void k_cells_assign(const float4* __restrict__ pos, int N,
                                                             double L, double qs, float inv_side,
                                                             int nc,
                                                             int* __restrict__ count,
                                                             int* __restrict__ cellof,
                                                             int* __restrict__ rank,
                                                             unsigned* __restrict__ flag)
{
  const int i = blockIdx.x * kCellBlock + threadIdx.x;
  if (i >= N) return;
  uint4 rec;
  bool bad;
  const int c = cell_record<WALL>(pos[i], i, L, qs, inv_side, nc, rec, bad);
  cellof[i] = c;
  rank[i] = atomicAdd(&count[c], 1);
  if (bad) atomicOr(flag, 1u);
}

// The same with the counts in LDS (ncell <= kCellLds): each block ranks its
// particles in a private histogram, then adds each cell's count to the
// cell's total with one global atomic, whose old value is the block's offset
// within the cell (bcnt[block * ncell + c]).  One atomic per (block, cell)
// instead of one per particle, which on a few hundred counters serialised
// (28 us at N = 16384 over 125 cells, against 3 us).
constexpr int kCellLds = 4096;
constexpr int kAssignPer = 4;  // particles per thread of the LDS-counted assignment
template <bool WALL>
__global__ __launch_bounds__(kCellBlock) void k_cells_assign_lds(const float4* __restrict__ pos, int N,
                                                                 double L, double qs,
                                                                 float inv_side, int nc,
                                                                 int ncell,
                                                                 int* __restrict__ count,
                                                                 int* __restrict__ bcnt,
                                                                 int* __restrict__ cellof,
                                                                 int* __restrict__ rank,
                                                                 unsigned* __restrict__ flag)
{
  __shared__ int h[kCellLds];
  for (int c = threadIdx.x; c < ncell; c += kCellBlock) h[c] = 0;
  __syncthreads();
  for (int k = 0; k < kAssignPer; ++k) {
    const int i = blockIdx.x * (kCellBlock * kAssignPer) + k * kCellBlock + (int)threadIdx.x;
    if (i >= N) break;
    uint4 rec;
    bool bad;
    const int c = cell_record<WALL>(pos[i], i, L, qs, inv_side, nc, rec, bad);
    cellof[i] = c;
    rank[i] = atomicAdd(&h[c], 1);
    if (bad) atomicOr(flag, 1u);
  }
  __syncthreads();
  // the block's place within each cell it holds particles of (the blocks'
  // order within a cell varies from pass to pass; the counts do not)
  for (int c = threadIdx.x; c < ncell; c += kCellBlock)
    if (h[c]) bcnt[(size_t)blockIdx.x * ncell + c] = atomicAdd(&count[c], h[c]);
}

// After k_cells_assign_lds, one launch over the same blocks: each block scans
// the cells' totals into their starts (block 0 also writes them to start) and
// scatters its own particles to start + the block's offset + rank.  The scan
// is repeated by every block (ncell reads each, from L2) instead of a
// separate one-block launch; the totals are cleared by k_rdf_cells_reduce.
template <bool WALL>
__global__ __launch_bounds__(kCellBlock) void k_cells_place(const float4* __restrict__ pos, int N,
                                                            double qs, int ncell,
                                                            const int* __restrict__ count,
                                                            const int* __restrict__ bcnt,
                                                            const int* __restrict__ cellof,
                                                            const int* __restrict__ rank,
                                                            int* __restrict__ start,
                                                            uint4* __restrict__ sorted)
{
  __shared__ int cs[kCellLds];
  __shared__ int wsum[kWaves];
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  int carry = 0;
  for (int base = 0; base < ncell; base += kCellBlock) {
    const int k = base + (int)threadIdx.x;
    const int tot = k < ncell ? count[k] : 0;
    int x = tot;  // inclusive scan inside the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int before = carry, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      before += w < wv ? wsum[w] : 0;
      total += wsum[w];
    }
    if (k < ncell) {
      const int ex = before + x - tot;
      cs[k] = ex;
      if (blockIdx.x == 0) start[k] = ex;
    }
    carry += total;
    __syncthreads();  // wsum is rewritten by the next chunk; cs is complete after the last
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) start[ncell] = carry;
  for (int u = 0; u < kAssignPer; ++u) {
    const int i = blockIdx.x * (kCellBlock * kAssignPer) + u * kCellBlock + (int)threadIdx.x;
    if (i >= N) break;
    const float4 p = pos[i];
    const int c = cellof[i];
    sorted[cs[c] + bcnt[(size_t)blockIdx.x * ncell + c] + rank[i]] =
        WALL ? make_uint4(__float_as_uint(p.x), __float_as_uint(p.y), __float_as_uint(p.z), (unsigned)i)
             : make_uint4(fixq(p.x, qs), fixq(p.y, qs), fixq(p.z, qs), (unsigned)i);
  }
}

// exclusive scan of count[0, n) into start[0, n], one block; count is
// cleared for the next pass
__global__ __launch_bounds__(1024) void k_cells_scan(int* __restrict__ count, int* __restrict__ start,
                                                     int n)
{
  // per chunk of 1024: an inclusive scan inside each wave (shuffles), the
  // 16 wave totals scanned by wave 0, then each value offset by its wave's
  // exclusive total and the chunks' running carry
  __shared__ int wsum[16];
  __shared__ int carry;
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int k = base + (int)threadIdx.x;
    const int v = k < n ? count[k] : 0;
    if (k < n) count[k] = 0;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (wv == 0) {
      int t = lane < 16 ? wsum[lane] : 0;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        const int y = __shfl_up(t, o, 64);
        if (lane >= o) t += y;
      }
      if (lane < 16) wsum[lane] = t;  // inclusive
    }
    __syncthreads();
    const int before = carry + (wv > 0 ? wsum[wv - 1] : 0);
    if (k < n) start[k] = before + x - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += wsum[15];
    __syncthreads();
  }
  if (threadIdx.x == 0) start[n] = carry;
}

// the particles to their cells' starts + their ranks (k_cells_assign's
// global counters)
template <bool WALL>
__global__ __launch_bounds__(kCellBlock) void k_cells_scatter(const float4* __restrict__ pos, int N,
                                                              double qs,
                                                              const int* __restrict__ cellof,
                                                              const int* __restrict__ rank,
                                                              const int* __restrict__ start,
                                                              uint4* __restrict__ sorted)
{
  const int i = blockIdx.x * kCellBlock + threadIdx.x;
  if (i >= N) return;
  const float4 p = pos[i];
  sorted[start[cellof[i]] + rank[i]] =
      WALL ? make_uint4(__float_as_uint(p.x), __float_as_uint(p.y), __float_as_uint(p.z), (unsigned)i)
           : make_uint4(fixq(p.x, qs), fixq(p.y, qs), fixq(p.z, qs), (unsigned)i);
}

// the 13 forward neighbours of a cell (dz > 0, or dz = 0 and dy > 0, or
// dz = dy = 0 and dx > 0), after the cell itself (offset 0)
__device__ __forceinline__ void shell_offset(int o, int& dx, int& dy, int& dz)
{
  // o = 1..13 enumerate the 26 neighbours' forward half in lexicographic
  // (dz, dy, dx) order
  const int k = o + 13;  // index of (dx, dy, dz) in [-1, 1]^3, 13 = the cell itself
  dx = k % 3 - 1;
  dy = (k / 3) % 3 - 1;
  dz = k / 9 - 1;
}

// One workgroup's pairs: the particles of cell blockIdx.x against those of
// its neighbour blockIdx.y (0..13: the cell itself, then 13 forward
// neighbours), counted into replica (blockIdx.x + ncell blockIdx.y) % kRdfReps.
// Returns (uniformly over the workgroup) with nothing counted when either cell
// is empty or the flag is set.
template <int RULE>
__device__ __forceinline__ void rdf_cells_pairs(
    const uint4* __restrict__ sorted, const int* __restrict__ start, const float4* __restrict__ pos,
    int nc, double L, int nbins, const float* __restrict__ tab, float est_scale, float t0, float t1,
    float t2, float t3, double qscale, const unsigned* __restrict__ flag,
    unsigned* __restrict__ reps, unsigned long long* __restrict__ other)
{
  constexpr bool TAB = RULE != LJ_RDF_RBINS;
  constexpr bool GPU = RULE == LJ_RDF_REF_GPU;
  if (*flag) return;  // coordinates beyond the bound's span: no count
  const int c = blockIdx.x, o = blockIdx.y;
  const int cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
  int dx, dy, dz;
  shell_offset(o, dx, dy, dz);
  const int c2 = (cx + dx + nc) % nc + nc * ((cy + dy + nc) % nc + nc * ((cz + dz + nc) % nc));
  const int i0 = start[c], i1 = start[c + 1];
  const int j0 = start[c2], j1 = start[c2 + 1];
  if (i0 == i1 || j0 == j1) return;  // uniform, before any barrier

  // dynamic LDS: T2 [nbins] when TAB: (T[k], T[k + 1]) S^2; T [nbins + 1]; the
  // per-wave histograms
  extern __shared__ float2 sh_cells[];
  float2* T2 = sh_cells;
  float* T = reinterpret_cast<float*>(T2 + (TAB ? nbins : 0));
  unsigned int* hs = reinterpret_cast<unsigned int*>(T + nbins + 1);
  __shared__ unsigned long long s_other;
  __shared__ uint4 sj[kJStage];
  for (int b = threadIdx.x; b < kWaves * nbins; b += kCellBlock) hs[b] = 0u;
  unsigned int* hw = hs + (int)(threadIdx.x >> 6) * nbins;
  const double S2 = qscale * qscale;
  for (int b = threadIdx.x; b <= nbins; b += kCellBlock) {
    const float t = tab[b];
    T[b] = t;
    if constexpr (TAB) {
      const float tq = (float)((double)t * S2);
      if (b < nbins) T2[b].x = tq;
      if (b > 0) T2[b - 1].y = tq;
    }
  }
  if (threadIdx.x == 0) s_other = 0ull;
  __syncthreads();

  const float thr[4] = {t0, t1, t2, t3};
  const float Tmax = T[nbins], Tlast = T[nbins - 1];
  RdfClass<RULE> cl;
  cl.init(3. * L, L, qscale, est_scale, nbins, Tlast, Tmax);
  // GPU rule: the clamp bin takes every pair not counted elsewhere; the pass
  // counts the pairs of the other bins (its histogram's total, taken at the
  // flush) and the excluded pairs r2 <= 1e-6 (this lane's, from the exact path)
  unsigned int nexcl = 0u;

  // The neighbour cell's particles are staged in LDS, kJStage at a time (so
  // that the walks below gather them from LDS, not global memory); each wave
  // takes its own chunks of 64 i (a cell of ~130 particles keeps three waves
  // busy; a wave without a chunk idles to the next barrier).
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  for (int jb = j0; jb < j1; jb += kJStage) {
    const int jn = min(kJStage, j1 - jb);
    __syncthreads();  // the previous chunk is no longer read
    for (int t = threadIdx.x; t < jn; t += kCellBlock) sj[t] = sorted[jb + t];
    __syncthreads();
    for (int ib = i0 + 64 * wave; ib < i1; ib += kCellBlock) {
      const int ii = ib + lane;
      const bool vi = ii < i1;
      const uint4 qi = sorted[vi ? ii : i1 - 1];
      // the lane's pairs: j in the chunk, and j > ii within the cell itself
      const int jlo = vi ? (o == 0 ? ii : jb - 1) : j1;
      // all j of the chunk classified 8 at a time (about 13% of an adjacent
      // cell pair's pairs lie in range: batches keep the table reads and
      // atomics of independent pairs in flight together); the rare undecided
      // pairs are walked through the reference arithmetic per run of 32
      for (int c0 = 0; c0 < jn; c0 += 32) {
        const int cn = min(32, jn - c0);
        unsigned int mask = 0u;
        constexpr int B = 8;
        for (int b0 = 0; b0 < cn; b0 += B) {
          int kk[B];
          bool dec[B], gone[B];
          if constexpr (TAB) {
            float lo[B], hi[B];
            float2 tt[B];
#pragma unroll
            for (int b = 0; b < B; ++b) {
              cl.measure(rdf_r2q(qi, sj[min(c0 + b0 + b, jn - 1)]), kk[b], lo[b], hi[b]);
              tt[b] = T2[kk[b]];
            }
#pragma unroll
            for (int b = 0; b < B; ++b) cl.decide(kk[b], lo[b], hi[b], tt[b], dec[b], gone[b]);
          } else {
#pragma unroll
            for (int b = 0; b < B; ++b)
              cl.rbins(rdf_r2q(qi, sj[min(c0 + b0 + b, jn - 1)]), kk[b], dec[b], gone[b]);
          }
#pragma unroll
          for (int b = 0; b < B; ++b) {
            const int t = c0 + b0 + b;
            const bool ok = (t < jn) & (jb + t > jlo);
            if (ok & dec[b]) atomicAdd(&hw[kk[b]], 1u);
            mask |= (ok & !dec[b] & !gone[b]) ? (1u << (b0 + b)) : 0u;
          }
        }
        while (mask) {
          const int t = c0 + __builtin_ctz(mask);
          mask &= mask - 1u;
          const int bin = rdf_exact_bin<RULE, true>(pos[qi.w], pos[sj[t].w], L, (float)L, thr, T,
                                                    nbins, Tmax, Tlast, est_scale);
          if constexpr (GPU) {
            if (bin != nbins - 1) {
              if (bin >= 0) atomicAdd(&hw[bin], 1u);
              else ++nexcl;
            }
          } else {
            if (bin >= 0) atomicAdd(&hw[bin], 1u);
          }
        }
      }
    }
  }
  __syncthreads();
  unsigned* rep = reps + (size_t)((blockIdx.x + gridDim.x * blockIdx.y) % kRdfReps) * nbins;
  unsigned long long tot = nexcl;  // GPU rule: this thread's share of the pass's counted pairs
  for (int b = threadIdx.x; b < nbins; b += kCellBlock) {
    unsigned int v = 0u;
    for (int w = 0; w < kWaves; ++w) v += hs[w * nbins + b];
    if (v) atomicAdd(&rep[b], v);
    tot += v;
  }
  if constexpr (GPU) {
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) tot += __shfl_xor(tot, s, 64);
    if (lane == 0 && tot) atomicAdd(&s_other, tot);
    __syncthreads();
    if (threadIdx.x == 0 && s_other)
      atomicAdd(&other[(blockIdx.x + gridDim.x * blockIdx.y) % kRdfReps], s_other);
  }
}

// one workgroup of kRdfReps threads per bin: hist[b] += 2 sum of the bin's
// replicas (cleared); workgroup 0 also adds the GPU rule's clamp bin,
// 2 (P - sum of other), clears other and the flag and sets result when the
// pass was void.  A workgroup that reads the flag after workgroup 0 cleared
// it adds only zeros (a void pass leaves every replica at zero).
template <int RULE>
__global__ __launch_bounds__(kCellBlock) void k_rdf_cells(
    const uint4* __restrict__ sorted, const int* __restrict__ start, const float4* __restrict__ pos,
    int nc, double L, int nbins, const float* __restrict__ tab, float est_scale, float t0, float t1,
    float t2, float t3, double qscale, const unsigned* __restrict__ flag,
    unsigned* __restrict__ reps, unsigned long long* __restrict__ other)
{
  rdf_cells_pairs<RULE>(sorted, start, pos, nc, L, nbins, tab, est_scale, t0, t1, t2, t3, qscale,
                        flag, reps, other);
}

// Reach 2: cells of side at least sqrt(X) / 2, so every pair below the far
// bound lies within two cells on each axis.  The candidate volume per
// particle is 125 cells of side s / 2 (a 5^3 block, 125 / 8 = 15.6 cells of
// side s) against 27 cells of side s for reach 1, 42% less.  One workgroup
// per cell: the cell's own particles and those of its 62 forward neighbours
// in [-2, 2]^3 are staged in LDS (self first, kJStage2 at a time), and the
// 256 lanes are spread over the cell's n particles and 256 / n phases of the
// staged j, so cells of any occupancy keep almost every lane busy.  The
// classification and counting are those of k_rdf_cells.
constexpr int kNb2 = 63;         // the cell itself, then its 62 forward neighbours
constexpr int kJStage2 = 1536;   // staged j (24 KiB)

// offset o = 0..62 of the reach-2 half shell: (dx, dy, dz) in [-2, 2]^3 at
// lexicographic (dz, dy, dx) index 62 + o (62 = the cell itself)
__device__ __forceinline__ void shell2_offset(int o, int& dx, int& dy, int& dz)
{
  const int k = o + 62;
  dx = k % 5 - 2;
  dy = (k / 5) % 5 - 2;
  dz = k / 25 - 2;
}

// WALL: a hard-wall box.  The records hold float coordinates, neighbours
// beyond the grid are left out (no wrap), and every candidate pair takes the
// reference arithmetic directly (r^2 and the bin from the threshold table):
// there is no image to approximate.
template <int RULE, bool WALL>
__global__ __launch_bounds__(kCellBlock) void k_rdf_cells2(
    const uint4* __restrict__ sorted, const int* __restrict__ start, const float4* __restrict__ pos,
    int nc, double L, int nbins, const float* __restrict__ tab, float est_scale, float t0, float t1,
    float t2, float t3, double qscale, const unsigned* __restrict__ flag,
    unsigned* __restrict__ reps, unsigned long long* __restrict__ other)
{
  constexpr bool TAB = RULE != LJ_RDF_RBINS;
  constexpr bool GPU = RULE == LJ_RDF_REF_GPU;
  if (*flag) return;  // coordinates beyond the bound's span: no count
  const int c = blockIdx.x;
  const int i0 = start[c], ni = start[c + 1] - i0;
  if (ni == 0) return;  // uniform, before any barrier
  const int cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);

  // dynamic LDS as k_rdf_cells
  extern __shared__ float2 sh_cells[];
  float2* T2 = sh_cells;
  float* T = reinterpret_cast<float*>(T2 + (TAB ? nbins : 0));
  unsigned int* hs = reinterpret_cast<unsigned int*>(T + nbins + 1);
  __shared__ unsigned long long s_other;
  __shared__ int s_cell[kNb2];
  __shared__ int s_off[kNb2 + 1];  // s_off[k]: staged index of neighbour k's first particle
  __shared__ uint4 sj[kJStage2];
  for (int b = threadIdx.x; b < kWaves * nbins; b += kCellBlock) hs[b] = 0u;
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  unsigned int* hw = hs + wave * nbins;
  const double S2 = qscale * qscale;
  for (int b = threadIdx.x; b <= nbins; b += kCellBlock) {
    const float t = tab[b];
    T[b] = t;
    if constexpr (TAB) {
      const float tq = (float)((double)t * S2);
      if (b < nbins) T2[b].x = tq;
      if (b > 0) T2[b - 1].y = tq;
    }
  }
  if (wave == 0) {  // the neighbours, their counts scanned across the wave
    int cnt = 0;
    if (lane < kNb2) {
      int dx, dy, dz;
      shell2_offset(lane, dx, dy, dz);
      int c2;
      bool in = true;
      if constexpr (WALL) {
        const int x = cx + dx, y = cy + dy, z = cz + dz;
        in = x >= 0 && x < nc && y >= 0 && y < nc && z >= 0 && z < nc;
        c2 = in ? x + nc * (y + nc * z) : c;
      } else {
        c2 = (cx + dx + nc) % nc + nc * ((cy + dy + nc) % nc + nc * ((cz + dz + nc) % nc));
      }
      s_cell[lane] = c2;
      cnt = in ? start[c2 + 1] - start[c2] : 0;  // a neighbour beyond a wall holds nothing
    }
    int x = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane < kNb2) s_off[lane + 1] = x;
    if (lane == 0) s_off[0] = 0;
  }
  if (threadIdx.x == 0) s_other = 0ull;
  __syncthreads();
  const int total = s_off[kNb2];
  const int nself = s_off[1];  // the cell's own particles, staged first

  const float thr[4] = {t0, t1, t2, t3};
  const float Tmax = T[nbins], Tlast = T[nbins - 1];
  RdfClass<RULE> cl;
  cl.init(3. * L, L, qscale, est_scale, nbins, Tlast, Tmax);
  unsigned int nexcl = 0u;  // GPU rule: excluded pairs r2 <= 1e-6 (this lane's)

  const int tid = (int)threadIdx.x;
  for (int ib = 0; ib < ni; ib += kCellBlock) {
    const int n = min(ni - ib, kCellBlock);
    const int P = kCellBlock / n;  // phases over the staged j
    const int li = tid % n, ph = tid / n;
    const bool act = ph < P;
    const int iloc = ib + li;  // the lane's particle, by its place in the cell
    const uint4 qi = sorted[i0 + iloc];
    const float4 pi = make_float4(__uint_as_float(qi.x), __uint_as_float(qi.y),
                                  __uint_as_float(qi.z), 0.f);  // WALL
    for (int jb = 0; jb < total; jb += kJStage2) {
      const int jn = min(kJStage2, total - jb);
      __syncthreads();  // the previous stage is no longer read
      for (int e = tid; e < jn; e += kCellBlock) {
        const int g = jb + e;
        int lo = 0, hi = kNb2;  // s_off[lo] <= g < s_off[hi]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_off[mid] <= g) lo = mid;
          else hi = mid;
        }
        sj[e] = sorted[start[s_cell[lo]] + (g - s_off[lo])];
      }
      __syncthreads();
      if (WALL && act) {
        // the reference arithmetic for every candidate, 8 at a time
        for (int tb = ph; tb < jn; tb += 8 * P) {
          constexpr int B = 8;
          float r2[B];
          int kk[B];
          bool ok[B];
#pragma unroll
          for (int b = 0; b < B; ++b) {
            const int t = tb + b * P, g = jb + t;
            const uint4 rj = sj[min(t, jn - 1)];
            r2[b] = rdf_wall_r2<RULE>(pi, make_float4(__uint_as_float(rj.x), __uint_as_float(rj.y),
                                                      __uint_as_float(rj.z), 0.f));
            ok[b] = (t < jn) & ((g >= nself) | (g > iloc));
            const float e = RULE == LJ_RDF_RBINS ? __builtin_amdgcn_sqrtf(r2[b]) * est_scale
                                                 : r2[b] * est_scale;
            kk[b] = r2[b] >= Tlast ? nbins - 1 : rdf_bin(r2[b], T, nbins, (int)e);
          }
#pragma unroll
          for (int b = 0; b < B; ++b) {
            bool count;
            if constexpr (GPU) {  // counted below the clamp bin; r2 <= 1e-6 excluded
              const bool in = r2[b] > 1e-6f;
              nexcl += (ok[b] & !in) ? 1u : 0u;
              count = ok[b] & in & (r2[b] < Tlast);
            } else {  // dropped beyond the last bin
              count = ok[b] & (r2[b] < Tmax);
            }
            if (count) atomicAdd(&hw[kk[b]], 1u);
          }
        }
      } else if (act) {
        // j = t0 + b P, 8 at a time; in the cell itself only j after the
        // lane's particle (each unordered pair once)
        for (int tb = ph; tb < jn; tb += 8 * P) {
          constexpr int B = 8;
          int kk[B];
          bool dec[B], gone[B];
          if constexpr (TAB) {
            float lo[B], hi[B];
            float2 tt[B];
#pragma unroll
            for (int b = 0; b < B; ++b) {
              cl.measure(rdf_r2q(qi, sj[min(tb + b * P, jn - 1)]), kk[b], lo[b], hi[b]);
              tt[b] = T2[kk[b]];
            }
#pragma unroll
            for (int b = 0; b < B; ++b) cl.decide(kk[b], lo[b], hi[b], tt[b], dec[b], gone[b]);
          } else {
#pragma unroll
            for (int b = 0; b < B; ++b)
              cl.rbins(rdf_r2q(qi, sj[min(tb + b * P, jn - 1)]), kk[b], dec[b], gone[b]);
          }
          unsigned int mask = 0u;
#pragma unroll
          for (int b = 0; b < B; ++b) {
            const int t = tb + b * P, g = jb + t;
            const bool ok = (t < jn) & ((g >= nself) | (g > iloc));
            if (ok & dec[b]) atomicAdd(&hw[kk[b]], 1u);
            mask |= (ok & !dec[b] & !gone[b]) ? (1u << b) : 0u;
          }
          while (mask) {
            const int t = tb + __builtin_ctz(mask) * P;
            mask &= mask - 1u;
            const int bin = rdf_exact_bin<RULE, true>(pos[qi.w], pos[sj[t].w], L, (float)L, thr, T,
                                                      nbins, Tmax, Tlast, est_scale);
            if constexpr (GPU) {
              if (bin != nbins - 1) {
                if (bin >= 0) atomicAdd(&hw[bin], 1u);
                else ++nexcl;
              }
            } else {
              if (bin >= 0) atomicAdd(&hw[bin], 1u);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  unsigned* rep = reps + (size_t)(blockIdx.x % kRdfReps) * nbins;
  unsigned long long tot = nexcl;  // GPU rule: this thread's share of the pass's counted pairs
  for (int b = threadIdx.x; b < nbins; b += kCellBlock) {
    unsigned int v = 0u;
    for (int w = 0; w < kWaves; ++w) v += hs[w * nbins + b];
    if (v) atomicAdd(&rep[b], v);
    tot += v;
  }
  if constexpr (GPU) {
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) tot += __shfl_xor(tot, s, 64);
    if (lane == 0 && tot) atomicAdd(&s_other, tot);
    __syncthreads();
    if (threadIdx.x == 0 && s_other) atomicAdd(&other[blockIdx.x % kRdfReps], s_other);
  }
}

// Sums the replicas into hist and clears them (one wave per bin); the GPU
// rule's clamp bin gets every pair the pass did not count elsewhere, 2 (P -
// other) as ordered pairs; clears the flag after noting it in result for the
// host.  (The same done by the pair kernel's last workgroup, behind a release
// count per workgroup, measured 67 us against 51 + 4.6: on this chip an
// agent-scope release writes back the XCD's L2.)
__global__ __launch_bounds__(kRdfReps) void k_rdf_cells_reduce(
    unsigned* __restrict__ reps, unsigned long long* __restrict__ other, int nbins, int gpu,
    unsigned long long P, unsigned* __restrict__ flag, unsigned* __restrict__ result,
    unsigned long long* __restrict__ hist, int* __restrict__ count, int ncell)
{
  static_assert(kRdfReps == 64, "one wave per bin");
  const int b = blockIdx.x, r = threadIdx.x;
  // the cells' totals of the LDS-counted sort, zero for the next pass
  for (int c = b * kRdfReps + r; c < ncell; c += gridDim.x * kRdfReps) count[c] = 0;
  const bool bad = *flag != 0u;
  unsigned long long s = reps[(size_t)r * nbins + b];
  reps[(size_t)r * nbins + b] = 0u;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (r == 0 && !bad && s) atomicAdd(&hist[b], 2ull * s);
  if (b == 0) {
    unsigned long long so = other[r];
    other[r] = 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) so += __shfl_xor(so, o, 64);
    if (r == 0) {
      if (!bad && gpu) atomicAdd(&hist[nbins - 1], 2ull * (P - so));
      if (bad) *result = 1u;
      *flag = 0u;
    }
  }
}

}  // namespace

namespace {

// sqrt(X) in fixed-point units: the side a reach-1 cell needs (0 for an
// invalid range)
double rdf_far_side(double L, double end)
{
  if (!(L > 0.) || !(end > 0.) || !std::isfinite(end)) return 0.;
  // the device's far bound at the guaranteed span A = 3 L (RdfClass::far_bound)
  const double qs = 4294967296. / L, S2 = qs * qs, u = 0x1p-24, A = 3. * L;
  const double Krd = 1.01 * 3.4641016151377546 * (2. * A + 3. * L) * u * qs;
  const double K2d = 1.01 * 16. * u;
  const double K0d = 1.01 * (3. * u * u * (2. * A + 5. * L) * (2. * A + 5. * L) * S2 + Krd * Krd);
  const double Tq = (double)(float)end * S2;
  const double dl2 = Krd * std::sqrt(2. * Tq) + K2d * 2. * Tq + K0d;
  // the float rounding of X up, and 1e-6 beside it: the cell side (in
  // fixed-point units) must square to at least X
  const double X = (Tq + 2. * dl2) * (1. + 0x1p-20) * (1. + 1e-6);
  return std::sqrt(X);
}

}  // namespace

// reach-1 cells per axis for a range ending at `end` (real r^2 units): at
// most kRdfMaxCells, or 0 when fewer than 3 fit
int rdf_cells_per_axis(double L, double end)
{
  const double side = rdf_far_side(L, end);
  if (!(side > 0.)) return 0;
  // larger cells than needed are as exact: at most kRdfMaxCells per axis
  const int nc = (int)std::min((double)kRdfMaxCells, std::floor(4294967296. / side));
  return nc >= 3 ? nc : 0;
}

// cells per axis and reach (1: side >= sqrt(X), the cell and 13 forward
// neighbours; 2: side >= sqrt(X) / 2, the cell and 62 forward neighbours) for
// a periodic range ending at r^2 = end, or 0 where reach 1 gives fewer than 3
// cells per axis.  The layout with the smaller candidate volume, 27 / nc1^3
// against 125 / nc2^3 of the box, is taken; reach 2 needs 5 or more cells per
// axis (distinct neighbours).  LJ_RDF_CELLS=1 / 2 holds the reach at 1 / 2
// (2 where it applies).
int rdf_cells_layout(double L, double end, int* reach)
{
  *reach = 1;
  const int nc1 = rdf_cells_per_axis(L, end);
  if (nc1 <= 0) return 0;
  const double side = rdf_far_side(L, end);
  const int nc2 = (int)std::min((double)kRdfMaxCells, std::floor(2. * 4294967296. / side));
  static const char* e = std::getenv("LJ_RDF_CELLS");
  if (nc2 >= 5) {
    const bool auto2 = 125. / ((double)nc2 * nc2 * nc2) < 27. / ((double)nc1 * nc1 * nc1);
    if ((e && e[0] == '2') || (!(e && e[0] == '1') && auto2)) {
      *reach = 2;
      return nc2;
    }
  }
  return nc1;
}

int rdf_cells_ensure(RdfCellsWs& w, int N, int ncell, int nbins, hipStream_t st, bool* grown)
{
  *grown = false;
  if (N <= w.cap_n && ncell <= w.cap_cells && nbins <= w.cap_bins && w.flag) return LJ_OK;
  *grown = w.flag != nullptr;  // arrays an earlier capture may hold
  // the arrays may be in use by earlier passes on the owner's stream (only
  // that stream: a device-wide wait can break another thread's capture)
  if (hipStreamSynchronize(st) != hipSuccess)
    return fail(LJ_ERR_DEVICE, "rdf_cells_ensure: synchronizing the stream failed");
  // grown to at least the old capacities, so that alternating layouts settle
  const int cn = std::max(N, w.cap_n), cc = std::max(ncell, w.cap_cells), cb = std::max(nbins, w.cap_bins);
  rdf_cells_free(w);
  LJ_TRY(dev_alloc(&w.count, cc));
  LJ_TRY(dev_alloc(&w.start, (size_t)cc + 1));
  LJ_TRY(dev_alloc(&w.cellof, cn));
  LJ_TRY(dev_alloc(&w.rank, cn));
  LJ_TRY(dev_alloc(&w.sorted, cn));
  LJ_TRY(dev_alloc(&w.reps, (size_t)kRdfReps * cb));
  LJ_TRY(dev_alloc(&w.other, kRdfReps));
  LJ_TRY(dev_alloc(&w.flag, 3));
  w.result = w.flag + 1;
  w.result_step = w.flag + 2;
  LJ_TRY(dev_alloc(&w.bcnt, kCellScanMax));
  w.cap_m = kCellScanMax;
  // the counts, replicas and flags start at zero; the passes clear them after use
  if (hipMemsetAsync(w.count, 0, sizeof(int) * cc, st) != hipSuccess ||
      hipMemsetAsync(w.reps, 0, sizeof(unsigned) * kRdfReps * cb, st) != hipSuccess ||
      hipMemsetAsync(w.other, 0, sizeof(unsigned long long) * kRdfReps, st) != hipSuccess ||
      hipMemsetAsync(w.flag, 0, sizeof(unsigned) * 3, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(LJ_ERR_DEVICE, "rdf_cells_ensure: clearing the workspace failed");
  w.cap_n = cn;
  w.cap_cells = cc;
  w.cap_bins = cb;
  return LJ_OK;
}

void rdf_cells_free(RdfCellsWs& w)
{
  dev_free(w.count);
  dev_free(w.start);
  dev_free(w.cellof);
  dev_free(w.rank);
  dev_free(w.sorted);
  dev_free(w.reps);
  dev_free(w.other);
  dev_free(w.flag);
  dev_free(w.bcnt);
  w.result = w.result_step = nullptr;
  w.cap_m = 0;
  w.cap_n = w.cap_cells = w.cap_bins = 0;
}

// the cell-list pass serves periodic boxes with at most 1024 bins and 3 or
// more cells per axis, given a workspace large enough (LJ_RDF_CELLS=0 turns
// it off)
// hard walls: reach-2 cells of side at least half the counted range's end
// (with a relative and an absolute margin over the float rounding of r^2
// and of the cell coordinates), 6 to 64 per axis, or 0 (fewer than 6: the
// neighbourhoods would cover most of the box)
int rdf_wall_cells_per_axis(double L, double end)
{
  if (!(L > 0.) || !(end > 0.) || !std::isfinite(end)) return 0;
  const double reach = std::sqrt((double)(float)end) * (1. + 1e-4) + 1e-4;
  const int nc = (int)std::min((double)kRdfMaxCells, std::floor(2. * L / reach));
  return nc >= 6 ? nc : 0;
}

int rdf_cells_grid(const RdfArgs& a, int* reach)
{
  *reach = 2;
  return a.periodic ? rdf_cells_layout(a.L, a.end, reach) : rdf_wall_cells_per_axis(a.L, a.end);
}

bool rdf_uses_cells(const RdfArgs& a)
{
  static const char* e = std::getenv("LJ_RDF_CELLS");
  if (e && e[0] == '0') return false;
  if (!a.cells || a.nbins > 1024) return false;
  int reach = 1;
  const int nc = rdf_cells_grid(a, &reach);
  return nc > 0 && a.cells->flag && a.N <= a.cells->cap_n && nc * nc * nc <= a.cells->cap_cells &&
         a.nbins <= a.cells->cap_bins;
}

template <bool WALL>
void launch_cells_sort(hipStream_t st, const RdfArgs& a, const RdfCellsWs& w, int nc, double qs,
                       float inv_side)
{
  const int ncell = nc * nc * nc;
  const int nb = (a.N + kCellBlock - 1) / kCellBlock;
  const int na = (a.N + kCellBlock * kAssignPer - 1) / (kCellBlock * kAssignPer);
  if (ncell <= kCellLds && (long)ncell * na <= (long)w.cap_m) {
    // per-block counts in LDS; the cells' starts and the blocks' offsets
    // found and the particles placed by the same blocks in a second launch
    hipLaunchKernelGGL(k_cells_assign_lds<WALL>, dim3(na), dim3(kCellBlock), 0, st, a.pos, a.N, a.L,
                       qs, inv_side, nc, ncell, w.count, w.bcnt, w.cellof, w.rank, w.flag);
    hipLaunchKernelGGL(k_cells_place<WALL>, dim3(na), dim3(kCellBlock), 0, st, a.pos, a.N, qs, ncell,
                       w.count, w.bcnt, w.cellof, w.rank, w.start, w.sorted);
  } else {
    hipLaunchKernelGGL(k_cells_assign<WALL>, dim3(nb), dim3(kCellBlock), 0, st, a.pos, a.N, a.L, qs,
                       inv_side, nc, w.count, w.cellof, w.rank, w.flag);
    hipLaunchKernelGGL(k_cells_scan, dim3(1), dim3(1024), 0, st, w.count, w.start, ncell);
    hipLaunchKernelGGL(k_cells_scatter<WALL>, dim3(nb), dim3(kCellBlock), 0, st, a.pos, a.N, qs,
                       w.cellof, w.rank, w.start, w.sorted);
  }
}

hipError_t launch_rdf_cells(hipStream_t st, const RdfArgs& a)
{
  const RdfCellsWs& w = *a.cells;
  int reach = 1;
  const int nc = rdf_cells_grid(a, &reach);
  const int ncell = nc * nc * nc;
  const double qs = 4294967296. / a.L;
  const bool wall = !a.periodic;
  if (wall) launch_cells_sort<true>(st, a, w, nc, qs, (float)(nc / a.L));
  else launch_cells_sort<false>(st, a, w, nc, qs, 0.f);
  const size_t lds = (a.rule != LJ_RDF_RBINS ? sizeof(float2) * a.nbins : 0) +
                     sizeof(float) * (a.nbins + 1) + sizeof(unsigned) * kWaves * a.nbins;
#define LJ_CELLS_ARGS w.sorted, w.start, a.pos, nc, a.L, a.nbins, a.tab, a.est_scale, a.thr[0], \
                      a.thr[1], a.thr[2], a.thr[3], qs, w.flag, w.reps, w.other
#define LJ_CELLS_LAUNCH(K, G, ...)                                                               \
  do {                                                                                           \
    if (a.rule == LJ_RDF_REF_CPU)                                                                \
      hipLaunchKernelGGL((K<LJ_RDF_REF_CPU __VA_ARGS__>), G, dim3(kCellBlock), lds, st,          \
                         LJ_CELLS_ARGS);                                                         \
    else if (a.rule == LJ_RDF_REF_GPU)                                                           \
      hipLaunchKernelGGL((K<LJ_RDF_REF_GPU __VA_ARGS__>), G, dim3(kCellBlock), lds, st,          \
                         LJ_CELLS_ARGS);                                                         \
    else                                                                                         \
      hipLaunchKernelGGL((K<LJ_RDF_RBINS __VA_ARGS__>), G, dim3(kCellBlock), lds, st,            \
                         LJ_CELLS_ARGS);                                                         \
  } while (0)
  if (wall) LJ_CELLS_LAUNCH(k_rdf_cells2, dim3(ncell), , true);
  else if (reach == 2) LJ_CELLS_LAUNCH(k_rdf_cells2, dim3(ncell), , false);
  else LJ_CELLS_LAUNCH(k_rdf_cells, dim3(ncell, 14));
#undef LJ_CELLS_LAUNCH
#undef LJ_CELLS_ARGS
  const unsigned long long P = (unsigned long long)a.N * (unsigned long long)(a.N - 1) / 2ull;
  hipLaunchKernelGGL(k_rdf_cells_reduce, dim3(a.nbins), dim3(kRdfReps), 0, st, w.reps, w.other,
                     a.nbins, a.rule == LJ_RDF_REF_GPU ? 1 : 0, P, w.flag,
                     a.sampling ? w.result_step : w.result, a.hist, w.count, ncell);
  return hipGetLastError();
}

}  // namespace lj
