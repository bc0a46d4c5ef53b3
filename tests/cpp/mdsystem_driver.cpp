// mdsystem_driver.cpp -- drives ljhip::MDSystem (include/lj_mdsystem.hpp) the
// way the reference's task programs drive their MDSystem (run-isotherm.cpp:
// configure, Integrate in a loop, read the public fields) and prints what a
// test compares against the oracle:
//   init U T K V P
//   step k t U T K V P      (one line per step)
//   rdfstep <256 counts>     NdNdr2 as the last Integrate left it (every step bins)
//   posstep <4N floats>      h_Pos after the last step
//   flucs momN momN2 momN3 momN4 momiters   after Fluctuations(0.3) on that
//                            state, Integrate(dt), Fluctuations(0.3) (each call
//                            also prints the reference's "Iteration:" line)
//   rdf <256 counts>
//   pos <4N floats>          (h_Pos after the last step)
//   vel0 / velo0             velocities, then initvelo(3.0, 0.05) of them
//   vel1 / velo1             after 10 more steps, then updatevelo()
//   maxwell T M(0) M(0.02) ... M(8)
//   cpu_mode_rejected 1
// usage: mdsystem_driver N rho T0 canonical boundary dt steps seed
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <iostream>

#include "lj_mdsystem.hpp"

static void print_floats(const char* tag, const float* a, int n)
{
  std::printf("%s", tag);
  for (int i = 0; i < n; ++i) std::printf(" %.9g", a[i]);
  std::printf("\n");
}

static void print_curve(const char* tag, const ljhip::MDSystem::Curve& c)
{
  std::printf("%s", tag);
  for (size_t i = 0; i < c.x.size(); ++i) std::printf(" %.17g %.17g", c.x[i], c.y[i]);
  std::printf("\n");
}

int main(int argc, char** argv)
{
  if (argc != 9) {
    std::fprintf(stderr, "usage: %s N rho T0 canonical boundary dt steps seed\n", argv[0]);
    return 2;
  }
  ljhip::MDSystem::MDSystemConfiguration cfg;
  cfg.N = std::atoi(argv[1]);
  cfg.rho = std::atof(argv[2]);
  cfg.T0 = std::atof(argv[3]);
  cfg.canonical = std::atoi(argv[4]) != 0;
  cfg.boundaryConditions = std::atoi(argv[5]);
  const double dt = std::atof(argv[6]);
  const int steps = std::atoi(argv[7]);
  cfg.seed = std::strtoull(argv[8], nullptr, 10);
  try {
    ljhip::MDSystem s(cfg);
    std::printf("init %.17g %.17g %.17g %.17g %.17g\n", s.U, s.T, s.K, s.V, s.P);
    // Pshear: 0 as on the reference's GPU branch, then the CPU path's value
    const double ps_gpu_branch = s.Pshear;
    std::printf("pshear %.17g %.17g\n", ps_gpu_branch, s.UpdateShearPressure());
    for (int k = 0; k < steps; ++k) {
      s.Integrate(dt);
      std::printf("step %d %.17g %.17g %.17g %.17g %.17g %.17g\n", k, s.t, s.U, s.T, s.K, s.V, s.P);
    }
    std::printf("rdfstep");
    for (int c : s.NdNdr2) std::printf(" %d", c);
    std::printf("\n");
    print_floats("posstep", s.h_Pos, 4 * cfg.N);
    std::fflush(stdout);
    s.Fluctuations(0.3);
    std::cout.flush();
    s.Integrate(dt);
    print_floats("posstep2", s.h_Pos, 4 * cfg.N);
    std::fflush(stdout);
    s.Fluctuations(0.3);
    std::cout.flush();
    std::printf("flucs %.17g %.17g %.17g %.17g %d\n", s.momN, s.momN2, s.momN3, s.momN4, s.momiters);
    s.CalculateForces();
    std::printf("rdf");
    for (int c : s.NdNdr2) std::printf(" %d", c);
    std::printf("\npos");
    for (int i = 0; i < 4 * cfg.N; ++i) std::printf(" %.9g", s.h_Pos[i]);
    std::printf("\n");
    // speed distribution: initvelo on the current velocities, 10 more steps,
    // updatevelo (MDSystem.cpp:651-694)
    print_floats("vel0", s.getArray(1), 4 * cfg.N);
    s.initvelo(3.0, 0.05);
    print_curve("velo0", s.getvelo());
    s.Integrate(dt, 10);
    print_floats("vel1", s.getArray(1), 4 * cfg.N);
    s.updatevelo();
    print_curve("velo1", s.getvelo());
    std::printf("maxwell %.17g", s.T);
    for (int k = 0; k <= 400; ++k) std::printf(" %.17g", s.Maxwell(0.02 * k));
    std::printf("\n");
    int rejected = 0;
    try {
      s.setHardwareMode(false);
    } catch (const std::runtime_error&) {
      rejected = 1;
    }
    std::printf("cpu_mode_rejected %d\n", rejected);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
