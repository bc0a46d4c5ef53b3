/*
 * lj_oracle_flucs.c -- CPU restatement of the reference's fluctuation
 * observables and time averages.
 *
 * TEST INFRASTRUCTURE ONLY (the test oracle).  Not linked into the product.
 * Reference: vlvovch/lennard-jones-cuda
 *   src/tasks/run-fluctuations/include/run-fluctuations-aux.h:128-278
 *     GetNSubsystem, GetNsubVz, GetNSubsystemBatch, GetNsubVzBatch
 *   src/library/MDSystem.cpp:708-724          MDSystem::Fluctuations (count, moments)
 *   src/tasks/auxiliary/time-average-aux.h:27-67   TimeAverage
 *   src/extra/sample-moments/NumberStatistics.h:128-134,168-301,563-617
 *     (moment sums, central moments, mean / variance / scaled-variance errors)
 * Build: -O2 -ffp-contract=off, no -ffast-math, as the reference's x86-64 build.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "lj_oracle.h"

/* the alpha ladder of the batch counters: alpha = step, then alpha += step
 * (accumulated, not k * step) while alpha < limit; limit = 1 - 1e-9 for
 * coordinates (run-fluctuations-aux.h:198), 1 + 1e-9 for velocities (:251) */
int orc_fl_ladder(double step, double limit, double* alphas, int cap)
{
  int n = 0;
  for (double alpha = step; alpha < limit; alpha += step) {
    if (alphas && n < cap) alphas[n] = alpha;
    ++n;
  }
  return n;
}

/* first index k with t[k] > v (std::upper_bound) */
static int upper_bound_d(const double* t, int n, double v)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (v < t[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

/* GetNSubsystemBatch (run-fluctuations-aux.h:191-241): cumulative particle
 * counts of the ladder; type 3 = centred cubes of side L alpha^(1/3) (open),
 * types 0..2 = slabs coordinate / L < alpha.  Returns the ladder size. */
int orc_fl_nsub_batch(int N, const float* pos4, double L, double alpha_step, int type, int* cnts,
                      int cap)
{
  const int m = orc_fl_ladder(alpha_step, 1. - 1.e-9, NULL, 0);
  double tLs[4096];
  if (m > cap || m > 4096) return -1;
  int k = 0;
  for (double alpha = alpha_step; alpha < 1. - 1.e-9; alpha += alpha_step) {
    if (type == 3) tLs[k] = L * pow(alpha, 1. / 3.);
    cnts[k++] = 0;
  }
  for (int i = 0; i < 4 * N; i += 4) {
    double x = pos4[i], y = pos4[i + 1], z = pos4[i + 2];
    if (type == 3) {
      double dx = (x - 0.5 * L), dy = (y - 0.5 * L), dz = (z - 0.5 * L);
      int tx = upper_bound_d(tLs, m, fabs(dx) / 0.5);
      int ty = upper_bound_d(tLs, m, fabs(dy) / 0.5);
      int tz = upper_bound_d(tLs, m, fabs(dz) / 0.5);
      int t = tx > ty ? tx : ty;
      t = t > tz ? t : tz;
      if (t < m) cnts[t]++;
    } else {
      double coord = type == 1 ? y : type == 2 ? z : x;
      coord /= L;
      int t = (int)(coord / alpha_step);
      /* tind < cnts.size() compares as unsigned: a negative index (a
       * particle left of the box) is skipped */
      if ((unsigned)t < (unsigned)m) cnts[t]++;
    }
  }
  for (int i = 1; i < m; ++i) cnts[i] += cnts[i - 1];
  return m;
}

/* GetNsubVzBatch (run-fluctuations-aux.h:246-278): cumulative counts of
 * |v_type| / vcut_max < alpha over the ladder alpha < 1 + 1e-9 */
int orc_fl_vz_batch(int N, const float* vel4, double vcut_max, double alpha_step, int type,
                    int* cnts, int cap)
{
  const int m = orc_fl_ladder(alpha_step, 1. + 1.e-9, NULL, 0);
  if (m > cap) return -1;
  for (int k = 0; k < m; ++k) cnts[k] = 0;
  for (int i = 0; i < 4 * N; i += 4) {
    double v = type == 1 ? (double)vel4[i + 1] : type == 2 ? (double)vel4[i + 2] : (double)vel4[i];
    v /= vcut_max;
    int t = (int)(fabs(v) / alpha_step);
    if ((unsigned)t < (unsigned)m) cnts[t]++;
  }
  for (int i = 1; i < m; ++i) cnts[i] += cnts[i - 1];
  return m;
}

/* GetNSubsystem (run-fluctuations-aux.h:128-160) */
int orc_fl_nsub(int N, const float* pos4, double L, double fraction, int type)
{
  double tL = L * pow(fraction, 1. / 3.);
  int ret = 0;
  for (int i = 0; i < 4 * N; i += 4) {
    double x = pos4[i], y = pos4[i + 1], z = pos4[i + 2];
    if (type == 3) {
      double dx = (x - 0.5 * L), dy = (y - 0.5 * L), dz = (z - 0.5 * L);
      if (fabs(dx) < 0.5 * tL && fabs(dy) < 0.5 * tL && fabs(dz) < 0.5 * tL) ret++;
    } else {
      double coord = type == 1 ? y : type == 2 ? z : x;
      if (coord < fraction * L) ret++;
    }
  }
  return ret;
}

/* GetNsubVz (run-fluctuations-aux.h:165-186) */
int orc_fl_vz(int N, const float* vel4, double vcut, int type)
{
  int ret = 0;
  for (int i = 0; i < 4 * N; i += 4) {
    double v = type == 1 ? (double)vel4[i + 1] : type == 2 ? (double)vel4[i + 2] : (double)vel4[i];
    if (fabs(v) < vcut) ret++;
  }
  return ret;
}

/* the count of MDSystem::Fluctuations (MDSystem.cpp:708-719): closed centred
 * cube, float coordinates against double bounds */
long long orc_fl_closed_cube(int N, const float* pos4, double L, double fraction)
{
  double tsz = L * pow(fraction, 1. / 3.);
  long long tN = 0;
  for (int i = 0; i < 4 * N; i += 4) {
    if (pos4[i] >= 0.5 * L - 0.5 * tsz && pos4[i] <= 0.5 * L + 0.5 * tsz &&
        pos4[i + 1] >= 0.5 * L - 0.5 * tsz && pos4[i + 1] <= 0.5 * L + 0.5 * tsz &&
        pos4[i + 2] >= 0.5 * L - 0.5 * tsz && pos4[i + 2] <= 0.5 * L + 0.5 * tsz)
      tN++;
  }
  return tN;
}

/* ---- TimeAverage over NumberStatistics ---- */

void orc_tavg_init(orc_tavg* a) { memset(a, 0, sizeof(*a)); }

/* TimeAverage::AddObservation (time-average-aux.h:38-48) with
 * NumberStatistics::UpdateMomentSumsWithObservation (NumberStatistics.h:608-617):
 * sums[k] += N^k by successive products, k = 0..4 (the orders the time
 * averages read; the higher sums of the reference's nmax = 16 feed nothing
 * they print) */
void orc_tavg_add(orc_tavg* a, double val)
{
  double tN = 1.;
  for (int k = 0; k < 5; ++k) {
    a->sums[k] += tN;
    tN *= val;
  }
  a->nobs++;
  if (a->iters > 0) a->corr += a->last * val;
  a->last = val;
  a->iters++;
}

static int64_t binom(int n, int k)
{
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

/* NumberStatistics::CalculateMoments (NumberStatistics.h:563-604) for orders
 * 0..4: moments m[k] = sums[k] / n, central mu[mm] as the binomial sum */
static void moments(const orc_tavg* a, double m[5], double mu[5])
{
  for (int i = 0; i < 5; ++i) m[i] = a->sums[i] / a->nobs;
  mu[0] = 1.;
  mu[1] = 0.;
  for (int mm = 2; mm < 5; ++mm) {
    mu[mm] = 0.;
    double tmpn = 1.;
    for (int i = 0; i <= mm; ++i) {
      mu[mm] += binom(mm, i) * m[mm - i] * tmpn * ((i & 1) ? -1. : 1.);
      tmpn *= m[1];
    }
  }
}

/* out[ORC_TAVG_NOUT]:
 *   0 mean            GetMean (NumberStatistics.h:288)
 *   1 variance        GetVariance (:294)
 *   2 s               TimeAverage::GetS (time-average-aux.h:50-58)
 *   3 mean error      TimeAverage::GetMeanError (:64-66) = GetMeanError * sqrt(s)
 *   4 scaled variance GetScaledVariance (:299) = kappa_2 / mean
 *   5 scaled variance error GetScaledVarianceError (:301, GetScaledCumulantError :271-285)
 *   6 plain mean error NumberStatistics::GetMeanError (:291) */
void orc_tavg_eval(const orc_tavg* a, double* out)
{
  double m[5], mu[5];
  moments(a, m, mu);
  const double nE = (double)a->nobs;
  const double mean = m[1];
  const double var = mu[2];
  const double mean_err = sqrt((m[2] - m[1] * m[1]) / nE);
  const double ncorr = a->corr / (a->iters - 1) - mean * mean;
  double s = 2. / log(var / ncorr);
  if (s < 0.) s = 1.;
  /* kappa_2 from central moments: the one partition {0,1} with multiplier 1 */
  double k2 = 0.;
  {
    double tmp = 1.;
    tmp *= mu[2];
    k2 += tmp;
  }
  const double c1 = mu[2], c2 = mean;
  /* GetCentralMomentsSampleCovariance(2, 2) (NumberStatistics.h:212-229) */
  double c1dev = 0.;
  c1dev += mu[4];
  c1dev -= mu[2] * mu[2];
  c1dev += 2 * 2 * mu[2] * mu[1] * mu[1];
  c1dev -= 2 * mu[1] * mu[3];
  c1dev -= 2 * mu[3] * mu[1];
  c1dev /= nE;
  const double c2dev = mean_err * mean_err;
  const double c1c2cov = (mu[3] - 2 * mu[2] * mu[1]) / nE;
  out[0] = mean;
  out[1] = var;
  out[2] = s;
  out[3] = mean_err * sqrt(s);
  out[4] = k2 / mean;
  out[5] = fabs(c1 / c2) * sqrt(c1dev / c1 / c1 + c2dev / c2 / c2 - 2. * c1c2cov / c1 / c2);
  out[6] = mean_err;
}
