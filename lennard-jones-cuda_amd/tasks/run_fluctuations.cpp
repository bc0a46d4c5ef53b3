// run_fluctuations.cpp -- the fluctuations task of the reference
// (src/tasks/run-fluctuations/run-fluctuations.cpp:10-202) on device-resident
// steps.
//
// One box of N particles (TVN at T*, or EVN at energy u* per particle) is
// equilibrated for teq; then every production step samples, on the device
// inside the captured step:
//   - particle counts of slabs x, y, z < alpha L and of centred cubes of
//     volume fraction alpha, over the ladder alpha = spacing, 2 spacing, ...
//     (CoordFlucsAverage, run-fluctuations-aux.h:335-426);
//   - counts of |v_z| < alpha vcut_max, vcut_max = 3 sqrt(T*)
//     (MomentumFlucsAverage, :429-528);
//   - the slab z < L/2 count behind the console's w column (:133-135);
//   - the RDF of the step's force pass (RDF_average, :296-332);
// and every print_every steps (1000, as there) the console row, the row of
// <prefix>.TimeDep.txt and the files .RDF.dat, .flucsX/Y/Z/Cube/Vz.dat are
// written with the reference's columns.  The initial state is the lattice
// with seeded Gaussian velocities, not the reference's clock-seeded one.
//
// usage: run_fluctuations [parameter file]
//   parameters as input/N400.Tst1.4.rhost0.05, plus: seed (1),
//   timestamp (1; 0 leaves the date out of the output prefix), print_every (1000)
#include <algorithm>

#include "task_common.hpp"

using namespace ljtask;
using std::cout;
using std::endl;
using std::setw;

namespace {

constexpr double kPi = 3.141592653589793238462643;  // MDSystem.cpp:7
constexpr double kVfactor = 3.0;                    // run-fluctuations-aux.h:432

// CoordFlucsAverage::PrintToFile (run-fluctuations-aux.h:386-425)
void print_coord_flucs(const std::string& fn, const std::vector<double>& alphas, const double* st)
{
  std::ofstream fout(fn);
  fout << setw(15) << "alpha" << " " << setw(15) << "<N>" << " " << setw(15) << "w[N]" << " ";
  fout << setw(15) << "w[N]/(1-alpha)" << " " << setw(15) << "error" << " " << setw(15) << "s"
       << " ";
  fout << endl;
  for (size_t ia = 0; ia < alphas.size(); ++ia) {
    TimeAverage a;
    a.SetFromDeviceStats(st + ia * LJ_FL_NSTAT);
    const double alpha = alphas[ia];
    const double s = a.GetS();
    const double w = a.GetScaledVariance();
    fout << setw(15) << alpha << " " << setw(15) << a.GetMean() << " " << setw(15) << w << " ";
    fout << setw(15) << w / (1. - alpha) << " ";
    fout << setw(15) << a.GetScaledVarianceError() * std::sqrt(s) / (1. - alpha) << " ";
    fout << setw(15) << s << " " << endl;
  }
}

// MomentumFlucsAverage::PrintToFile (:485-527)
void print_vel_flucs(const std::string& fn, const std::vector<double>& vcuts, const double* st,
                     int totN)
{
  std::ofstream fout(fn);
  fout << setw(15) << "vcut" << " " << setw(15) << "<N>" << " " << setw(15) << "alpha" << " ";
  fout << setw(15) << "w[N]" << " " << setw(15) << "w[N]/(1-alpha)" << " " << setw(15) << "error"
       << " " << setw(15) << "s" << " ";
  fout << endl;
  for (size_t ia = 0; ia < vcuts.size(); ++ia) {
    TimeAverage a;
    a.SetFromDeviceStats(st + ia * LJ_FL_NSTAT);
    const double Nav = a.GetMean();
    const double s = a.GetS();
    const double w = a.GetScaledVariance();
    const double alpha = Nav / totN;
    fout << setw(15) << vcuts[ia] << " " << setw(15) << Nav << " " << setw(15) << alpha << " ";
    fout << setw(15) << w << " " << setw(15) << w / (1. - alpha) << " ";
    fout << setw(15) << a.GetScaledVarianceError() * std::sqrt(s) / (1. - alpha) << " ";
    fout << setw(15) << s << " " << endl;
  }
}

// RDF_average::PrintToFile (:316-331) of the time-averaged MDSystem::RDF
// (MDSystem.cpp:633-649) of the summed step histograms
void print_rdf(const std::string& fn, const std::vector<uint64_t>& sum, int64_t nsteps, int N,
               double L, float dr2)
{
  std::ofstream fout(fn);
  if (fout.is_open()) fout << setw(15) << "r*" << " " << setw(15) << "G(r)" << " ";
  fout << endl;
  const double n0 = N / L / L / L;
  for (size_t ir = 0; ir < sum.size(); ++ir) {
    const double r = std::sqrt((ir + 0.5) * dr2);
    const double counts = nsteps > 0 ? (double)sum[ir] / (double)nsteps : 0.;
    const double g = counts / dr2 / 2. / kPi / r / n0 / static_cast<double>(N);
    fout << setw(15) << r << " " << setw(15) << g << " " << endl;
  }
}

}  // namespace

int main(int argc, char* argv[])
{
  Parameters params;
  params.output_prefix = "run";
  params.parameters = {{"N", 400},       {"T*", 1.4},     {"u*", 1.708},       {"rho*", 0.60},
                       {"teq", 10.},     {"tfin", 1000.}, {"dt*", 0.004},      {"canonical", 1},
                       {"subvolume_spacing", 0.05},       {"useCUDA", 1}};
  if (argc > 1) {
    cout << "Reading parameters from file " << std::string(argv[1]) << endl;
    params.ReadParametersFromFile(std::string(argv[1]));
  }
  {
    // GetFullPrefix (run-fluctuations-aux.h:101-119)
    std::stringstream ss;
    const std::string ts = params.timestamp();
    if (!ts.empty()) ss << "." << ts;
    ss << ".N" << params.parameters["N"];
    if (params.iget("canonical"))
      ss << ".Tst" << params.parameters["T*"];
    else
      ss << ".ust" << params.parameters["u*"];
    ss << ".rhost" << params.parameters["rho*"];
    params.output_prefix += ss.str();
  }
  require_device(params);
  params.OutputParameters("fluctuations");

  const int N = (int)params.iget("N");
  const double rhost = params.get("rho*"), Tst = params.get("T*"), ust = params.get("u*");
  const double dt = params.get("dt*"), teq = params.get("teq"), tfin = params.get("tfin");
  const int canonical = (int)params.iget("canonical");
  const double dalpha = params.get("subvolume_spacing");
  const uint64_t seed = params.has("seed") ? (uint64_t)params.iget("seed") : 1;
  const long every = params.has("print_every") ? std::max(1L, params.iget("print_every")) : 1000;
  if (N <= 0 || !(dt > 0.) || !(dalpha > 0.)) die("N, dt* and subvolume_spacing must be positive");

  lj_config c;
  lj_config_init(&c);
  c.N = N;
  c.T0 = Tst;
  c.rho = rhost;
  c.canonical = canonical ? LJ_TVN : LJ_EVN;
  c.boundary = LJ_BC_PERIODIC;
  lj_ctx* ctx = nullptr;
  check(lj_create(&c, &ctx), "lj_create");
  check(lj_sample_initial_state(ctx, seed), "lj_sample_initial_state");
  if (!canonical) check(lj_renormalize_to_energy(ctx, ust), "lj_renormalize_to_energy");  // :64-66
  const double L = lj_box_length(ctx);
  const float dr2 = lj_rdf_dr2_default(N);

  // the counters: x, y, z slabs and cubes over the ladder, |v_z| windows, and
  // the z < L/2 slab of the console's w column
  const double vmax = std::sqrt(Tst) * kVfactor;  // :462
  const lj_fl_ladder lad[6] = {{LJ_FL_SLAB_X, 0, dalpha, 0.}, {LJ_FL_SLAB_Y, 0, dalpha, 0.},
                               {LJ_FL_SLAB_Z, 0, dalpha, 0.}, {LJ_FL_CUBE, 0, dalpha, 0.},
                               {LJ_FL_VEL_Z, 0, dalpha, vmax}, {LJ_FL_SLAB_Z, 1, 0.5, 0.}};
  check(lj_fl_configure(ctx, lad, 6), "lj_fl_configure");
  int E = 0;
  check(lj_fl_num_entries(ctx, &E), "lj_fl_num_entries");
  // alphas as CoordFlucsAverage and MomentumFlucsAverage build them (:351, :449)
  std::vector<double> alphas, vcuts;
  for (double alpha = dalpha; alpha <= 1. - 1.e-9; alpha += dalpha) alphas.push_back(alpha);
  for (double alpha = dalpha; alpha <= 1. + 1.e-9; alpha += dalpha)
    vcuts.push_back(std::sqrt(Tst) * alpha * kVfactor);
  const int mc = (int)alphas.size(), mv = (int)vcuts.size();
  if (E != 4 * mc + mv + 1) die("counter ladder sizes differ from the reference's alpha loops");

  double t = 0.;
  const long neq = count_steps(t, teq, dt);  // :82-85
  for (long done = 0; done < neq;) {
    const long n = std::min(10000L, neq - done);
    check(lj_integrate(ctx, dt, (int)n, nullptr), "lj_integrate");
    done += n;
  }

  cout << setw(15) << "t*" << " " << setw(15) << "u*" << " " << setw(15) << "T*" << " ";
  cout << setw(15) << "Z" << " " << setw(15) << "<u*>" << " " << setw(15) << "<T*>" << " ";
  cout << setw(15) << "<Z>" << " " << setw(15) << "<w>/(1-x)" << " " << endl;
  std::ofstream fout(params.output_prefix + std::string(".TimeDep.txt"));
  fout << setw(15) << "t*" << " " << setw(15) << "u*" << " " << setw(15) << "T*" << " ";
  fout << setw(15) << "Z" << " " << setw(15) << "<u*>" << " " << setw(15) << "<T*>" << " ";
  fout << setw(15) << "<Z>" << " " << setw(15) << "<wx>/(1-x)" << " ";
  fout << setw(15) << "<wy>/(1-x)" << " " << setw(15) << "<wz>/(1-x)" << " ";
  fout << setw(15) << "<vx>" << " " << setw(15) << "<vy>" << " " << setw(15) << "<vz>" << " ";
  fout << endl;

  // production: every step samples the counters and the RDF on the device
  check(lj_reset_averaging(ctx), "lj_reset_averaging");  // :117
  lj_step_sampling smp;
  smp.rdf = LJ_SAMPLE_RDF_SUM;
  smp.rdf_rule = LJ_RDF_REF_GPU;  // NdNdr2 as the reference's GPU path bins it
  smp.rdf_nbins = 256;
  smp.rdf_dr2 = dr2;
  smp.rdf_rmax = 0.;
  smp.fluctuations = 1;
  check(lj_set_step_sampling(ctx, &smp), "lj_set_step_sampling");
  std::vector<double> st((size_t)E * LJ_FL_NSTAT);
  std::vector<uint64_t> rsum(256);
  std::vector<float> vel((size_t)4 * N);
  long totIters = 0;
  while (t < tfin || tfin < 0.) {
    // steps up to the next output row (or the end of the run)
    double tt = t;
    long n = 0;
    while ((tt < tfin || tfin < 0.) && n < every - totIters % every) {
      tt += dt;
      ++n;
    }
    check(lj_integrate(ctx, dt, (int)n, nullptr), "lj_integrate");
    t = tt;
    totIters += n;
    if (totIters % every != 0) continue;
    lj_obs o;
    check(lj_get_observables(ctx, &o), "lj_get_observables");
    check(lj_fl_stats(ctx, st.data()), "lj_fl_stats");
    int64_t nrdf = 0;
    check(lj_step_rdf_counts(ctx, rsum.data(), &nrdf, 0), "lj_step_rdf_counts");
    check(lj_get_array(ctx, LJ_ARRAY_VEL, vel.data()), "lj_get_array");
    const double* half = st.data() + (size_t)(E - 1) * LJ_FL_NSTAT;  // the z < L/2 slab
    const double Nav = half[1] / totIters, N2av = half[2] / totIters;
    const double Tav = o.av_T_tot / o.av_iters;
    cout << setw(15) << t << " " << setw(15) << o.U / N << " " << setw(15) << o.T << " ";
    cout << setw(15) << o.P / (rhost * o.T) << " ";
    cout << setw(15) << o.av_U_tot / o.av_iters / N << " " << setw(15) << Tav << " ";
    cout << setw(15) << o.av_p_tot / o.av_iters / (rhost * Tav) << " ";
    cout << setw(15) << (N2av - Nav * Nav) / Nav / (1. - 0.5) << " " << endl;

    fout << setw(15) << t << " " << setw(15) << o.U / N << " " << setw(15) << o.T << " ";
    fout << setw(15) << o.P / (rhost * o.T) << " ";
    fout << setw(15) << o.av_U_tot / o.av_iters / N << " " << setw(15) << Tav << " ";
    fout << setw(15) << o.av_p_tot / o.av_iters / (rhost * Tav) << " ";
    for (int ic = 0; ic < 3; ++ic) {  // :164-181, the middle alpha of each slab ladder
      const int ti = mc / 2;
      const double* e = st.data() + (size_t)(ic * mc + ti) * LJ_FL_NSTAT;
      const double tNav = e[1] / e[7], tN2av = e[2] / e[7];
      const double tw = (tN2av - tNav * tNav) / tNav;
      fout << setw(15) << tw / (1. - alphas[ti]) << " ";
    }
    double av[3] = {0., 0., 0.};  // GetAvVel (:282-292)
    for (int i = 0; i < 4 * N; i += 4) {
      av[0] += vel[i];
      av[1] += vel[i + 1];
      av[2] += vel[i + 2];
    }
    for (int k = 0; k < 3; ++k) fout << setw(15) << av[k] / N << " ";
    fout << endl;
    fout.flush();

    print_rdf(params.output_prefix + ".RDF.dat", rsum, nrdf, N, L, dr2);
    const char* names[4] = {".flucsX.dat", ".flucsY.dat", ".flucsZ.dat", ".flucsCube.dat"};
    for (int ic = 0; ic < 4; ++ic)
      print_coord_flucs(params.output_prefix + names[ic], alphas,
                        st.data() + (size_t)ic * mc * LJ_FL_NSTAT);
    print_vel_flucs(params.output_prefix + ".flucsVz.dat", vcuts,
                    st.data() + (size_t)4 * mc * LJ_FL_NSTAT, N);
  }
  lj_destroy(ctx);
  return 0;
}
