// task_common.hpp -- what the task programs share: the reference's parameter
// file format and run-parameter printout, the time average with its error
// estimate, and the step-count arithmetic of the reference's time loops.
//
// Parameter files (src/tasks/run-*/include/run-*-aux.h, ReadParametersFromFile):
// whitespace-separated `name value` pairs; a token starting with '#' makes the
// rest of its line a comment; `output_prefix` takes a word, every other name a
// number.  Unknown names are kept, so files may carry these programs' extra
// parameters (seed, timestamp, print_every, batched).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "lj_hip.h"

namespace ljtask {

struct Parameters {
  std::string output_prefix;
  std::map<std::string, double> parameters;

  // run-isotherm-aux.h:53-84 / run-fluctuations-aux.h:54-85
  void ReadParametersFromFile(const std::string& filename)
  {
    std::ifstream fin(filename);
    if (!fin.is_open()) {
      std::cout << "Cannot open parameters file!" << "\n";
      std::cout.flush();
      return;
    }
    std::cout << "Reading input parameters from file " << filename << "\n";
    std::string name;
    while (fin >> name) {
      if (name.empty() || name[0] == '#') {
        fin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
        continue;
      }
      std::cout << "Reading input parameter " << name << " = ";
      if (name == "output_prefix") {
        fin >> output_prefix;
        std::cout << output_prefix << std::endl;
        continue;
      }
      double value = 0.;
      fin >> value;
      std::cout << value << std::endl;
      parameters[name] = value;
    }
    std::cout.flush();
  }

  void OutputParameters(const char* title)
  {
    const int tabsize = 25;
    std::cout << "Lennard-Jones Molecular Dynamics " << title << " run parameter list:" << "\n";
    for (auto& el : parameters)
      std::cout << std::setw(tabsize) << el.first << " = " << el.second << "\n";
    std::cout << std::setw(tabsize) << "output_prefix" << " = " << output_prefix << "\n";
    std::cout << std::endl;
  }

  double get(const char* name) { return parameters[name]; }
  long iget(const char* name) { return std::lround(parameters[name]); }
  bool has(const char* name) const { return parameters.count(name) != 0; }

  // the date-time part of the reference's output prefix (GetFullPrefix);
  // `timestamp 0` leaves it out, for reproducible file names
  std::string timestamp()
  {
    if (has("timestamp") && iget("timestamp") == 0) return "";
    auto t = std::time(nullptr);
    auto tm = *std::localtime(&t);
    std::stringstream ss;
    ss << std::put_time(&tm, "%d-%m-%Y-T%H-%M-%S");
    return ss.str();
  }
};

// The time average of one observable with its statistical error from the
// statistical inefficiency s (time-average-aux.h:27-67, after Allen and
// Tildesley pp. 194-195), over the moment sums of the observations: sum x^k,
// k = 0..4 (successive products, as NumberStatistics keeps them), the
// lag-one product sum and the last value.  The device's fluctuation counters
// accumulate exactly these sums (lj_fl_stats); host observables add them here.
class TimeAverage {
 public:
  void AddObservation(double x)
  {
    double p = 1.;
    for (int k = 0; k < 5; ++k) {
      sums_[k] += p;
      p *= x;
    }
    ++n_;
    if (iters_ > 0) corr_ += last_ * x;
    last_ = x;
    ++iters_;
  }

  // the sums of one entry of lj_fl_stats / lj_batch_fl_stats (LJ_FL_NSTAT)
  void SetFromDeviceStats(const double* s)
  {
    for (int k = 0; k < 5; ++k) sums_[k] = s[k];
    corr_ = s[5];
    last_ = s[6];
    n_ = (int64_t)s[7];
    iters_ = (int)s[7];
  }

  int64_t count() const { return n_; }
  double GetMean() const { return moment(1); }
  // central moment of order 2..4 from the ordinary moments (binomial sum in
  // the order of NumberStatistics::CalculateMoments)
  double GetCentral(int order) const
  {
    double mu = 0.;
    double pw = 1.;
    const double m1 = moment(1);
    for (int i = 0; i <= order; ++i) {
      mu += (double)binom(order, i) * moment(order - i) * pw * ((i & 1) ? -1. : 1.);
      pw *= m1;
    }
    return mu;
  }
  double GetVariance() const { return GetCentral(2); }
  // plain standard error of the mean, sqrt((<x^2> - <x>^2) / n)
  double GetPlainMeanError() const
  {
    return std::sqrt((moment(2) - moment(1) * moment(1)) / (double)n_);
  }
  // statistical inefficiency s = 2 / ln(var / lag-one covariance), 1 when negative
  double GetS() const
  {
    const double mean = GetMean();
    const double cov = corr_ / (iters_ - 1) - mean * mean;
    double s = 2. / std::log(GetVariance() / cov);
    if (s < 0.) s = 1.;
    return s;
  }
  double GetMeanError() const { return GetPlainMeanError() * std::sqrt(GetS()); }
  // kappa_2 / <x>
  double GetScaledVariance() const { return GetCentral(2) / GetMean(); }
  // error of kappa_2 / <x> from the sample covariances of mu_2 and <x>
  double GetScaledVarianceError() const
  {
    const double nE = (double)n_;
    const double mu1 = 0., mu2 = GetCentral(2), mu3 = GetCentral(3), mu4 = GetCentral(4);
    double c1dev = 0.;
    c1dev += mu4;
    c1dev -= mu2 * mu2;
    c1dev += 2 * 2 * mu2 * mu1 * mu1;
    c1dev -= 2 * mu1 * mu3;
    c1dev -= 2 * mu3 * mu1;
    c1dev /= nE;
    const double me = GetPlainMeanError();
    const double c2dev = me * me;
    const double cov = (mu3 - 2 * mu2 * mu1) / nE;
    const double c1 = mu2, c2 = GetMean();
    return std::fabs(c1 / c2) * std::sqrt(c1dev / c1 / c1 + c2dev / c2 / c2 - 2. * cov / c1 / c2);
  }

 private:
  double sums_[5] = {0., 0., 0., 0., 0.};
  int64_t n_ = 0;
  double corr_ = 0., last_ = 0.;
  int iters_ = 0;

  double moment(int k) const { return sums_[k] / n_; }
  static int64_t binom(int n, int k)
  {
    int64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
  }
};

// iterations of the reference's loop `while (t < tend) { step; t += dt; }`
// starting from t (advanced in place, accumulated as there)
inline long count_steps(double& t, double tend, double dt)
{
  long n = 0;
  while (t < tend) {
    t += dt;
    ++n;
  }
  return n;
}

inline void die(const std::string& msg)
{
  std::cerr << "error: " << msg << std::endl;
  std::exit(1);
}

inline void check(int rc, const char* what)
{
  if (rc != LJ_OK) die(std::string(what) + ": " + lj_last_error());
}

// the reference checks for a device and falls back to its CPU path
// (run-*.cpp:20-31); this library has no CPU path, so no device is an error,
// and `useCUDA 0` still runs on the device
inline void require_device(Parameters& p)
{
  if (lj_device_count() == 0)
    die("no HIP device found; this library has no CPU path (the reference's would run on the CPU)");
  if (p.has("useCUDA") && p.iget("useCUDA") == 0)
    std::cout << "useCUDA 0: this library runs on the GPU only; running on the GPU\n";
}

}  // namespace ljtask
