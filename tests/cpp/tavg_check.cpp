// tavg_check.cpp -- prints the statistics of the task programs' TimeAverage
// (lennard-jones-cuda_amd/tasks/task_common.hpp) for the values read from
// stdin, one per line: mean, variance, s, mean error, scaled variance, scaled
// variance error, plain mean error (%.17g), and the same after a round trip
// through the device-statistics layout (SetFromDeviceStats).
#include <cstdio>
#include <iostream>

#include "task_common.hpp"

int main()
{
  ljtask::TimeAverage a;
  double x;
  double sums[5] = {0, 0, 0, 0, 0}, corr = 0., last = 0.;
  long n = 0;
  while (std::cin >> x) {
    a.AddObservation(x);
    double p = 1.;
    for (int k = 0; k < 5; ++k) {
      sums[k] += p;
      p *= x;
    }
    if (n > 0) corr += last * x;
    last = x;
    ++n;
  }
  const double st[8] = {sums[0], sums[1], sums[2], sums[3], sums[4], corr, last, (double)n};
  ljtask::TimeAverage b;
  b.SetFromDeviceStats(st);
  for (const ljtask::TimeAverage* t : {&a, &b})
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", t->GetMean(), t->GetVariance(),
                t->GetS(), t->GetMeanError(), t->GetScaledVariance(), t->GetScaledVarianceError(),
                t->GetPlainMeanError());
  return 0;
}
