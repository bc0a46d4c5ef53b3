// CPU check of the RDF kernel's lookup tables (lennard-jones-cuda_amd/csrc/
// lj_host.cpp: cpu_minimage_thresholds, gpu_minimage_thresholds,
// rdf_bin_thresholds) and of the kernel's table arithmetic (lj_kernels.hip:
// mi_cpu, mi_gpu, rdf_bin), against the direct arithmetic of the reference's
// rules, restated here:
//   CPU image index   k = fast_round((float)(r / L))      (MDSystem.cpp:274, :732-739)
//   GPU image index   k = lround(d / (float)L), float     (MDSystem.cu:31-33)
//   CPU bin           floor((double)r2 / (double)dr2), dropped from nbins on (MDSystem.cpp:282)
//   GPU bin           floor(r2 / dr2) in float, clamped into the last bin (MDSystem.cu:39-41)
//   r-uniform bins    k with edges2[k] <= r2 < edges2[k + 1], edges2[k] = fl32((k rmax / nbins)^2)
// Inputs: uniform separations and squared distances plus values within a few
// ulps of every table entry, where a wrong threshold would show.  Prints one
// line per check with its mismatch count; the exit status is the total.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "lj_internal.h"

using namespace lj;

namespace {

float ulp_step(float x, int d)
{
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += (uint32_t)d;
  std::memcpy(&x, &u, 4);
  return x;
}

int fast_round(float x) { return x > 0 ? (int)(x + 0.5f) : (int)(x - 0.5f); }

// the kernel's threshold count (mi_cpu, mi_gpu)
int table_index(float r, const float (&t)[4])
{
  int k = (r >= t[0]) + (r >= t[1]) + (r >= t[2]) + (r >= t[3]);
  k -= (r <= -t[0]) + (r <= -t[1]) + (r <= -t[2]) + (r <= -t[3]);
  return k;
}

// the kernel's bin from an estimate (rdf_bin)
int table_bin(float r2, const float* T, int nbins, int est)
{
  int k = std::min(std::max(est, 0), nbins - 1);
  k -= T[k] > r2 ? 1 : 0;
  k += T[k + 1] <= r2 ? 1 : 0;
  return k;
}

long check_images(std::mt19937_64& rng, bool gpu)
{
  long bad = 0;
  for (double L : {20.0, 26.812329187, 13.406164593, 75.88, 7.5, 3.17}) {
    float t[4];
    if (gpu) gpu_minimage_thresholds(L, t);
    else cpu_minimage_thresholds(L, t);
    const float Lf = (float)L;
    std::uniform_real_distribution<float> U(-2.4f * Lf, 2.4f * Lf);
    for (int n = 0; n < 1000000; ++n) {
      float r = U(rng);
      if (n % 2) r = ulp_step(t[rng() % 4], (int)(rng() % 7) - 3) * ((rng() & 1) ? -1.f : 1.f);
      const int direct = gpu ? (int)std::lround(r / Lf) : fast_round((float)(r / L));
      if (std::abs(direct) > 4) continue;  // the tables cover |k| <= 4
      bad += table_index(r, t) != direct;
    }
  }
  return bad;
}

long check_bins(std::mt19937_64& rng, int rule)
{
  long bad = 0;
  for (float dr2 : {0.1f, 0.2f, 0.0517f, 0.31f})
    for (double rmax : {10.0, 13.406164593, 37.94})
      for (int nb : {3, 256, 500, 2000}) {
        std::vector<float> T(nb + 1);
        rdf_bin_thresholds(rule, nb, dr2, rmax, T.data());
        const float es = rule == LJ_RDF_RBINS ? (float)(nb / rmax) : (float)(1.0 / dr2);
        const float top = rule == LJ_RDF_RBINS ? T[nb] : 1.3f * nb * dr2;
        const int kmax = rule == LJ_RDF_REF_GPU ? nb - 1 : nb;  // T[nb] = inf: clamped, never dropped
        std::uniform_real_distribution<float> U(0.f, top);
        for (int n = 0; n < 60000; ++n) {
          float r2 = U(rng);
          if (n % 2) r2 = ulp_step(T[1 + (int)(rng() % kmax)], (int)(rng() % 5) - 2);
          if (!(r2 >= 0.f) || !std::isfinite(r2)) continue;
          // the direct rule: bin, or -1 when dropped
          int direct;
          if (rule == LJ_RDF_REF_CPU) {
            const double q = std::floor((double)r2 / (double)dr2);
            direct = q >= nb ? -1 : (int)q;
          } else if (rule == LJ_RDF_REF_GPU) {
            direct = std::min((int)std::floor(r2 / dr2), nb - 1);
          } else {
            direct = -1;
            for (int k = 0; k < nb; ++k)
              if (T[k] <= r2 && r2 < T[k + 1]) direct = k;
          }
          // the kernel: dropped beyond T[nbins] (not for the GPU rule), the last
          // bin counted apart from T[nbins - 1] on, else the corrected estimate
          int kern;
          if (!(r2 < T[nb])) {
            kern = -1;
          } else if (r2 >= T[nb - 1]) {
            kern = nb - 1;
          } else {
            for (int d = -1; d <= 1; ++d) {  // the hardware root within one ulp either way
              const float e = rule == LJ_RDF_RBINS ? ulp_step(std::sqrt(r2), d) * es : r2 * es;
              bad += table_bin(r2, T.data(), nb, (int)e) != direct;
            }
            continue;
          }
          bad += kern != direct;
        }
      }
  return bad;
}

}  // namespace

int main()
{
  std::mt19937_64 rng(11);
  const long a = check_images(rng, false), b = check_images(rng, true);
  const long c = check_bins(rng, LJ_RDF_REF_CPU), d = check_bins(rng, LJ_RDF_REF_GPU),
             e = check_bins(rng, LJ_RDF_RBINS);
  std::printf("cpu image index: %ld mismatches\ngpu image index: %ld mismatches\n", a, b);
  std::printf("cpu bins: %ld mismatches\ngpu bins: %ld mismatches\nr-uniform bins: %ld mismatches\n", c,
              d, e);
  return (int)std::min(a + b + c + d + e, 100L);
}
