// lj_fluct.hip -- fluctuation observables on the device (lj_fluct.h): one
// pass over the particles bins each particle into every ladder with the
// reference's double arithmetic; one thread per (box, ladder) then forms the
// cumulative counts and adds them to the accumulated moments.
//
// Counting: each workgroup covers kBlock particles of one box and keeps a
// private histogram of the box's entries in LDS (integer atomics), then adds
// its non-zero bins to the box's histogram in global memory (integer atomics:
// order-independent, so the counts are exact and deterministic).  The
// finaliser reads and clears that histogram, so it is zero again for the next
// sample without a memset node in the step graphs.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lj_fluct.h"
#include "lj_util.h"

namespace lj {

namespace {

constexpr int kMaxEntries = 4096;  // entries per box (LDS histogram of 16 KiB)

__device__ __forceinline__ float axis_of(const float4& p, int a)
{
  return a == 0 ? p.x : a == 1 ? p.y : p.z;
}

// first k with t[k] > v (std::upper_bound over the ascending table)
__device__ __forceinline__ int upper_bound_dev(const double* __restrict__ t, int m, double v)
{
  int lo = 0, hi = m;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v < t[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// (int)q of the reference's index arithmetic where the index can be counted:
// q in (-1, m) truncates to 0..m-1; anything else (negative indices, which
// the reference's unsigned size comparison skips, too large, NaN) gives -1
__device__ __forceinline__ int trunc_index(double q, int m)
{
  if (!(q > -1.0 && q < (double)m)) return -1;
  return (int)q;
}

// entry index of one particle in one ladder (-1: counted by no entry)
__device__ int fl_index(const FlLadderDev& l, const double* __restrict__ t, const float4& p,
                        const float4& v)
{
#pragma clang fp contract(off)
  switch (l.kind) {
    case FL_TABLE_CUBE: {
      // run-fluctuations-aux.h:210-219: dx = x - 0.5 L (lo holds 0.5 L),
      // index = the largest upper_bound(|dx| / 0.5) over the axes
      const int ix = upper_bound_dev(t, l.m, fabs((double)p.x - l.lo) / 0.5);
      const int iy = upper_bound_dev(t, l.m, fabs((double)p.y - l.lo) / 0.5);
      const int iz = upper_bound_dev(t, l.m, fabs((double)p.z - l.lo) / 0.5);
      return max(ix, max(iy, iz));
    }
    case FL_TABLE_COORD:  // run-fluctuations-aux.h:148-155 with one window
      return upper_bound_dev(t, l.m, (double)axis_of(p, l.axis));
    case FL_TABLE_VABS:   // run-fluctuations-aux.h:175-182
      return upper_bound_dev(t, l.m, fabs((double)axis_of(v, l.axis)));
    case FL_DIV_COORD: {  // run-fluctuations-aux.h:224-233
      double c = (double)axis_of(p, l.axis);
      c /= l.scale;
      return trunc_index(c / l.step, l.m);
    }
    case FL_DIV_VEL: {    // run-fluctuations-aux.h:262-271
      double w = (double)axis_of(v, l.axis);
      w /= l.scale;
      return trunc_index(fabs(w) / l.step, l.m);
    }
    case FL_CLOSED_CUBE: {  // MDSystem.cpp:714-718
      const double x = p.x, y = p.y, z = p.z;
      const bool in = x >= l.lo && x <= l.hi && y >= l.lo && y <= l.hi && z >= l.lo && z <= l.hi;
      return in ? 0 : -1;
    }
    default:
      return -1;
  }
}

__global__ __launch_bounds__(kBlock) void k_fl_count(const float4* __restrict__ pos,
                                                     const float4* __restrict__ vel, int N,
                                                     const BoxParam* __restrict__ box,
                                                     const int2* __restrict__ pblock,
                                                     const FlLadderDev* __restrict__ lad,
                                                     const double* __restrict__ tab,
                                                     unsigned* __restrict__ counts, int nl, int E)
{
  extern __shared__ unsigned hist[];
  int b = 0, start = (int)blockIdx.x * kBlock, n = N, off = 0;
  if (box) {
    const int2 pb = pblock[blockIdx.x];
    b = pb.x;
    start = pb.y;
    n = box[b].N;
    off = box[b].off;
  }
  for (int e = threadIdx.x; e < E; e += kBlock) hist[e] = 0u;
  __syncthreads();
  const int i = start + (int)threadIdx.x;
  if (i < n) {
    const float4 p = pos[off + i];
    const float4 v = vel[off + i];
    const FlLadderDev* L = lad + (size_t)b * nl;
    const double* t = tab + (size_t)b * E;
    for (int l = 0; l < nl; ++l) {
      const FlLadderDev ld = L[l];
      const int k = fl_index(ld, t + ld.off, p, v);
      if ((unsigned)k < (unsigned)ld.m) atomicAdd(&hist[ld.off + k], 1u);
    }
  }
  __syncthreads();
  unsigned* c = counts + (size_t)b * E;
  for (int e = threadIdx.x; e < E; e += kBlock) {
    const unsigned h = hist[e];
    if (h) atomicAdd(&c[e], h);
  }
}

// one thread per (box, ladder): cumulative counts (run-fluctuations-aux.h:237-238),
// the last sample's counts, and the statistics of each entry in sample order:
// sum n^k by successive products (NumberStatistics.h:608-617) and the
// lag-one product sum of TimeAverage::AddObservation (time-average-aux.h:38-48)
__global__ __launch_bounds__(64) void k_fl_finalize(unsigned* __restrict__ counts,
                                                    long long* __restrict__ last,
                                                    double* __restrict__ stats,
                                                    const FlLadderDev* __restrict__ lad, int nbox,
                                                    int nl, int E)
{
#pragma clang fp contract(off)
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (g >= nbox * nl) return;
  const int b = g / nl;
  const FlLadderDev l = lad[g];
  long long c = 0;
  for (int k = 0; k < l.m; ++k) {
    const size_t e = (size_t)b * E + l.off + k;
    c += counts[e];
    counts[e] = 0u;
    last[e] = c;
    const double v = (double)c;
    double* s = stats + e * kFlStat;
    double tN = 1.;
    for (int j = FL_S0; j <= FL_S4; ++j) {
      s[j] += tN;
      tN *= v;
    }
    if (s[FL_ITERS] > 0.) s[FL_CORR] += s[FL_LAST] * v;
    s[FL_LAST] = v;
    s[FL_ITERS] += 1.;
  }
}

__global__ void k_fl_zero(double* __restrict__ stats, size_t ns, unsigned* __restrict__ counts,
                          long long* __restrict__ last, size_t ne)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ns) stats[i] = 0.;
  if (i < ne) {
    counts[i] = 0u;
    last[i] = 0;
  }
}

// the reference's ladder loop (accumulated alpha) below limit
std::vector<double> ladder(double step, double limit)
{
  std::vector<double> a;
  for (double alpha = step; alpha < limit; alpha += step) {
    a.push_back(alpha);
    if (a.size() > (size_t)kMaxEntries) break;
  }
  return a;
}

}  // namespace

void fl_free(FlSampler& s)
{
  dev_free(s.lad);
  dev_free(s.tab);
  dev_free(s.counts);
  dev_free(s.last);
  dev_free(s.stats);
  s.nl = s.E = s.nbox = s.maxm = 0;
  s.spec.clear();
  s.values.clear();
}

int fl_configure(FlSampler& s, hipStream_t st, int nbox, const double* L, const lj_fl_ladder* spec,
                 int nl)
{
  if (nl < 0 || (nl > 0 && !spec)) return fail(LJ_ERR_ARG, "lj_fl_configure: bad ladder list");
  if (nl > 64) return fail(LJ_ERR_ARG, "lj_fl_configure: at most 64 ladders (got %d)", nl);
  // entries of each ladder (box independent)
  std::vector<int> m(nl), off(nl);
  std::vector<std::vector<double>> vals(nl);
  int E = 0;
  for (int l = 0; l < nl; ++l) {
    const lj_fl_ladder& d = spec[l];
    const bool vel = d.type >= LJ_FL_VEL_X && d.type <= LJ_FL_VEL_Z;
    if (d.type < LJ_FL_SLAB_X || d.type > LJ_FL_CUBE_CLOSED)
      return fail(LJ_ERR_ARG, "lj_fl_configure: ladder %d has unknown type %d", l, d.type);
    if (d.type == LJ_FL_CUBE_CLOSED && !d.single)
      return fail(LJ_ERR_ARG, "lj_fl_configure: the closed cube (ladder %d) takes one fraction", l);
    if (vel && !(d.vmax > 0.) )
      return fail(LJ_ERR_ARG, "lj_fl_configure: ladder %d needs vmax > 0", l);
    if (d.single) {
      if (!vel && !(d.alpha > 0.))
        return fail(LJ_ERR_ARG, "lj_fl_configure: ladder %d needs a positive fraction", l);
      vals[l] = {vel ? d.vmax : d.alpha};
    } else {
      if (!(d.alpha > 0.) || !(d.alpha <= 1.))
        return fail(LJ_ERR_ARG, "lj_fl_configure: ladder %d step must be in (0, 1]", l);
      vals[l] = ladder(d.alpha, vel ? 1. + 1.e-9 : 1. - 1.e-9);
      if (vals[l].empty()) return fail(LJ_ERR_ARG, "lj_fl_configure: ladder %d is empty", l);
    }
    m[l] = (int)vals[l].size();
    off[l] = E;
    E += m[l];
    if (E > kMaxEntries)
      return fail(LJ_ERR_ARG, "lj_fl_configure: more than %d entries", kMaxEntries);
  }
  // per-box ladders and tables, with the reference's double arithmetic
  std::vector<FlLadderDev> hl((size_t)nbox * nl);
  std::vector<double> ht((size_t)nbox * std::max(E, 1), 0.);
  for (int b = 0; b < nbox; ++b) {
    for (int l = 0; l < nl; ++l) {
      const lj_fl_ladder& d = spec[l];
      FlLadderDev& g = hl[(size_t)b * nl + l];
      std::memset(&g, 0, sizeof(g));
      g.m = m[l];
      g.off = off[l];
      double* t = ht.data() + (size_t)b * E + off[l];
      const double Lb = L[b];
      if (d.type == LJ_FL_CUBE) {
        g.kind = FL_TABLE_CUBE;
        g.lo = 0.5 * Lb;
        for (int k = 0; k < m[l]; ++k) t[k] = Lb * std::pow(vals[l][k], 1. / 3.);  // :200, :132
      } else if (d.type <= LJ_FL_SLAB_Z) {
        g.axis = d.type - LJ_FL_SLAB_X;
        if (d.single) {
          g.kind = FL_TABLE_COORD;
          t[0] = d.alpha * Lb;  // coord < fraction * L, :154
        } else {
          g.kind = FL_DIV_COORD;
          g.step = d.alpha;
          g.scale = Lb;
        }
      } else if (d.type <= LJ_FL_VEL_Z) {
        g.axis = d.type - LJ_FL_VEL_X;
        if (d.single) {
          g.kind = FL_TABLE_VABS;
          t[0] = d.vmax;  // |v| < vcut, :181
        } else {
          g.kind = FL_DIV_VEL;
          g.step = d.alpha;
          g.scale = d.vmax;
        }
      } else {
        g.kind = FL_CLOSED_CUBE;
        const double tsz = Lb * std::pow(d.alpha, 1. / 3.);  // MDSystem.cpp:711
        g.lo = 0.5 * Lb - 0.5 * tsz;
        g.hi = 0.5 * Lb + 0.5 * tsz;
      }
    }
  }
  LJ_HIP(hipStreamSynchronize(st));
  fl_free(s);
  if (nl == 0) return LJ_OK;
  const size_t ne = (size_t)nbox * E;
  LJ_TRY(dev_alloc(&s.lad, hl.size()));
  LJ_TRY(dev_alloc(&s.tab, ne));
  LJ_TRY(dev_alloc(&s.counts, ne));
  LJ_TRY(dev_alloc(&s.last, ne));
  LJ_TRY(dev_alloc(&s.stats, ne * kFlStat));
  LJ_HIP(hipMemcpyAsync(s.lad, hl.data(), sizeof(FlLadderDev) * hl.size(), hipMemcpyHostToDevice, st));
  LJ_HIP(hipMemcpyAsync(s.tab, ht.data(), sizeof(double) * ne, hipMemcpyHostToDevice, st));
  LJ_HIP(hipStreamSynchronize(st));
  s.nbox = nbox;
  s.nl = nl;
  s.E = E;
  s.maxm = *std::max_element(m.begin(), m.end());
  s.spec.assign(spec, spec + nl);
  s.values.clear();
  for (int l = 0; l < nl; ++l) s.values.insert(s.values.end(), vals[l].begin(), vals[l].end());
  LJ_HIP(launch_fl_reset(st, s));
  LJ_HIP(hipStreamSynchronize(st));
  return LJ_OK;
}

hipError_t launch_fl_reset(hipStream_t st, const FlSampler& s)
{
  if (!s.active()) return hipSuccess;
  const size_t ne = (size_t)s.nbox * s.E;
  const size_t ns = ne * kFlStat;
  hipLaunchKernelGGL(k_fl_zero, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, s.stats, ns,
                     s.counts, s.last, ne);
  return hipGetLastError();
}

hipError_t launch_fl_sample(hipStream_t st, const FlSampler& s, const float4* pos, const float4* vel,
                            int N, const BoxParam* box, const int2* pblock, int nblocks)
{
  if (!s.active()) return hipSuccess;
  const int nb = box ? nblocks : (N + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_fl_count, dim3(nb), dim3(kBlock), sizeof(unsigned) * s.E, st, pos, vel, N,
                     box, pblock, s.lad, s.tab, s.counts, s.nl, s.E);
  const int nt = s.nbox * s.nl;
  hipLaunchKernelGGL(k_fl_finalize, dim3((nt + 63) / 64), dim3(64), 0, st, s.counts, s.last,
                     s.stats, s.lad, s.nbox, s.nl, s.E);
  return hipGetLastError();
}

}  // namespace lj
