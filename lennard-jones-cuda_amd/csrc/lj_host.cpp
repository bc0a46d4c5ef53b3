// lj_host.cpp -- host-side helpers of libljhip.so: initial states of the
// benchmark configurations, box rules and the exact binning tables of the RDF.
#include <cmath>
#include <limits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "lj_internal.h"

namespace lj {

namespace {
constexpr double kPi = 3.141592653589793238462643;

// SplitMix64 (Steele, Lea, Flood 2014): one 64-bit draw per call
inline uint64_t splitmix_next(uint64_t& state)
{
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline double uniform01(uint64_t& state)  // [0, 1) with 53 random bits
{
  return (double)(splitmix_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

// fast_round of the reference (MDSystem.cpp:732-739)
inline int ref_fast_round(float x) { return x > 0 ? (int)(x + 0.5f) : (int)(x - 0.5f); }

// the reference's image index of a float separation r (MDSystem.cpp:274)
inline int ref_image_index(float r, double L) { return ref_fast_round((float)(r / L)); }

inline uint32_t fbits(float f)
{
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float bitsf(uint32_t u)
{
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
}  // namespace

double box_length(int N, double rho) { return std::pow(N / rho, 1. / 3.); }  // MDSystem.cpp:70

float rdf_dr2_rule(int N)  // MDSystem.cpp:93-95
{
  const double a = 0.2 * std::sqrt(100. / N);
  float dr2 = (float)(a > 0.05 ? a : 0.05);
  if (250 * dr2 < 25.0) dr2 = (float)(25.0 / 250);
  return dr2;
}

void lattice_positions(int N, double L, float* pos4)  // MDSystem.cpp:150-168
{
  const int n = (int)std::ceil(std::pow((double)N, 1. / 3.));
  const double dL = L / n;
  for (int k = 0; k < N; ++k) {
    const int ix = k % n, iy = (k / n) % n, iz = k / (n * n);
    pos4[4 * k + 0] = (float)((ix + 0.5) * dL);
    pos4[4 * k + 1] = (float)((iy + 0.5) * dL);
    pos4[4 * k + 2] = (float)((iz + 0.5) * dL);
    pos4[4 * k + 3] = (float)(L / 150.f);
  }
}

namespace {
// CorrectTotalMomentum (MDSystem.cpp:183-202), then RenormalizeVelocities(true)
// (MDSystem.cpp:375-385, KineticTemperature :361-373), on float velocities
void momentum_and_temperature(int N, double T0, float* vel4)
{
  double px = 0., py = 0., pz = 0., mass = 0.;
  for (int k = 0; k < N; ++k) {
    px += vel4[4 * k];
    py += vel4[4 * k + 1];
    pz += vel4[4 * k + 2];
    mass += 1.;
  }
  for (int k = 0; k < N; ++k) {
    vel4[4 * k] = (float)(vel4[4 * k] + -px / mass);
    vel4[4 * k + 1] = (float)(vel4[4 * k + 1] + -py / mass);
    vel4[4 * k + 2] = (float)(vel4[4 * k + 2] + -pz / mass);
  }
  double tk = 0.;
  for (int k = 0; k < N; ++k) {
    const float* v = vel4 + 4 * k;
    const float s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    tk += s;
  }
  tk *= 1. / 3. / N;
  for (int k = 0; k < 3 * N; ++k) {
    float& v = vel4[4 * (k / 3) + k % 3];
    v = (float)(v * std::sqrt(T0 / tk));
  }
}
}  // namespace

void gaussian_velocities(int N, double T0, uint64_t seed, float* vel4)
{
  // component m = 3k + c of particle k; Box-Muller pairs in m order
  uint64_t st = seed;
  const double sT = std::sqrt(T0);
  const int M = 3 * N;
  for (int m = 0; m < M; m += 2) {
    const double u1 = 1.0 - uniform01(st);
    const double u2 = uniform01(st);
    const double rad = std::sqrt(-2.0 * std::log(u1));
    vel4[4 * (m / 3) + m % 3] = (float)(sT * (rad * std::cos(2.0 * kPi * u2)));
    if (m + 1 < M) vel4[4 * ((m + 1) / 3) + (m + 1) % 3] = (float)(sT * (rad * std::sin(2.0 * kPi * u2)));
  }
  for (int k = 0; k < N; ++k) vel4[4 * k + 3] = 0.f;
  momentum_and_temperature(N, T0, vel4);
}

void maxwell_velocities(int N, double T0, uint64_t seed, float* vel4)
{
  // MDSystem.cpp:36-53, 170-177: the speed v = sqrt(2 T) t with t = -ln(alpha),
  // alpha uniform in (0, 1) accepted when eta < p(alpha), eta uniform in
  // [0, 1.2) (the generator's probmax, MDSystem.h:51), p(alpha) = t^2 exp(-t^2)
  // / alpha -- t then has density proportional to t^2 exp(-t^2), the Maxwell
  // speed distribution at temperature T; the direction has cos(theta) uniform
  // in [-1, 1) and phi uniform in [0, 2 pi).  As in the reference the speed
  // sampler has a generator of its own (`sp`) beside the one of the
  // directions (`dr`); both are SplitMix64 streams of the seed.
  uint64_t sp = seed, dr = seed ^ 0x5DEECE66DA3B9F1Full;
  const double s2T = std::sqrt(2. * T0);
  for (int k = 0; k < N; ++k) {
    double t = 0.;
    for (;;) {
      const double alpha = ((double)(splitmix_next(sp) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      const double eta = uniform01(sp) * 1.2;
      const double tl = -std::log(alpha);
      if (eta < tl * tl * std::exp(-tl * tl) / alpha) {
        t = tl;
        break;
      }
    }
    const double v = s2T * t;
    const double cth = 2. * uniform01(dr) - 1.;
    const double ph = 2. * kPi * uniform01(dr);
    const double sth = std::sqrt(1. - cth * cth);
    vel4[4 * k] = (float)(v * sth * std::cos(ph));
    vel4[4 * k + 1] = (float)(v * sth * std::sin(ph));
    vel4[4 * k + 2] = (float)(v * cth);
    vel4[4 * k + 3] = 0.f;
  }
  momentum_and_temperature(N, T0, vel4);
}

void cpu_minimage_thresholds(double L, float thr[4])
{
  // k(x) = fast_round((float)(x/L)) is non-decreasing in x; bisect on the
  // ordered bit patterns of positive floats for the first x with k(x) >= m.
  for (int m = 1; m <= 4; ++m) {
    uint32_t lo = fbits(0.0f);
    uint32_t hi = fbits((float)((m + 1) * L));
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (ref_image_index(bitsf(mid), L) >= m) hi = mid;
      else lo = mid;
    }
    thr[m - 1] = bitsf(hi);
  }
}

void rdf_edges2(int nbins, double rmax, float* edges2)
{
  const double dr = rmax / nbins;
  for (int k = 0; k <= nbins; ++k) {
    const double e = (double)k * dr;
    edges2[k] = (float)(e * e);
  }
}

namespace {
// the least float x >= 0 with pred(x), for a predicate non-decreasing in x
// (bisection on the ordered bit patterns of non-negative floats)
template <class P>
float first_float(float guess, P pred)
{
  if (pred(0.f)) return 0.f;
  float h = guess > 0.f ? guess : 1.f;
  while (!pred(h) && h < 3.0e38f) h *= 2.f;
  if (!pred(h)) return std::numeric_limits<float>::infinity();
  uint32_t lo = fbits(0.f), hi = fbits(h);
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (pred(bitsf(mid))) hi = mid;
    else lo = mid;
  }
  return bitsf(hi);
}
}  // namespace


void gpu_minimage_thresholds(double L, float thr[4])
{
  // k(x) = lround(x / L) in float with a correctly rounded division
  // (MDSystem.cu:31-33): non-decreasing, odd in x
  const float Lf = (float)L;
  for (int m = 1; m <= 4; ++m)
    thr[m - 1] = first_float((float)(m + 1) * Lf, [&](float x) { return std::round(x / Lf) >= (float)m; });
}

void rdf_bin_thresholds(int rule, int nbins, float dr2, double rmax, float* T)
{
  if (rule == LJ_RDF_RBINS) {
    rdf_edges2(nbins, rmax, T);
    return;
  }
  T[0] = 0.f;
  const double d2 = dr2;
  for (int k = 1; k <= nbins; ++k) {
    if (rule == LJ_RDF_REF_CPU)  // floor((double)r2 / (double)dr2) (MDSystem.cpp:282)
      T[k] = first_float((float)((k + 1) * d2),
                         [&](float x) { return std::floor((double)x / d2) >= (double)k; });
    else  // floor of the float quotient (MDSystem.cu:39)
      T[k] = first_float((float)(k + 1) * dr2, [&](float x) { return std::floor(x / dr2) >= (float)k; });
  }
  if (rule == LJ_RDF_REF_GPU) T[nbins] = std::numeric_limits<float>::infinity();  // clamped, never dropped
}

}  // namespace lj
