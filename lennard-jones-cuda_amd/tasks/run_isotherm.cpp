// run_isotherm.cpp -- the isotherm task of the reference
// (src/tasks/run-isotherm/run-isotherm.cpp:10-170) on device-resident steps.
//
// For every density rho = rho*_min, += drho*, up to rho*_max (accumulated as
// there), a box of N particles at T* in TVN dynamics is equilibrated for teq,
// then u* = U/N and p* = P are time-averaged over the production run up to
// tfin, observed before each step (:109-133).  The results file
// <prefix>.dat has the reference's columns; the isothermal compressibility
// w = T / (dp/drho) comes from neighbouring densities (:147-161).
//
// The densities are independent boxes.  By default they run together as one
// batched ensemble (lj_batch: every kernel of a step is one launch over all
// boxes, the steps replayed from captured graphs, the per-step U and P of
// every box copied back in chunks); `batched 0` runs them one after another
// as single-box contexts.  The initial state is the lattice with seeded
// Gaussian velocities (box k: seed + k), not the reference's clock-seeded one.
//
// usage: run_isotherm [parameter file]
//   parameters as input/N400.Tst1.4.isotherm, plus: seed (1), batched (1),
//   timestamp (1; 0 leaves the date out of the output prefix), print_every (1000)
#include <algorithm>
#include <memory>

#include "task_common.hpp"

using namespace ljtask;
using std::cout;
using std::endl;
using std::setw;

namespace {

struct Density {
  double rho = 0.;
  TimeAverage uav, Pav;
  double T_end = 0.;
  std::ostringstream rows;  // the console rows of this density
};

// observation i of a density: the state before production step i (:110-112)
void observe(Density& d, int N, double U, double P, double T, double t, long iter, long every,
             double dt, double T0)
{
  d.uav.AddObservation(U / N);
  d.Pav.AddObservation(P);
  if (iter % every != 0) return;
  std::ostringstream& o = d.rows;
  o << setw(15) << d.rho << " " << setw(15) << t << " " << setw(15) << T << " ";
  o << setw(15) << d.uav.GetMean() << " " << setw(15) << d.uav.GetMeanError() << " ";
  o << setw(15) << d.uav.GetS() * dt << " ";
  o << setw(15) << d.Pav.GetMean() << " " << setw(15) << d.Pav.GetMeanError() << " ";
  o << setw(15) << d.Pav.GetS() * dt << " ";
  o << setw(15) << d.Pav.GetMean() / d.rho / T0 << " ";
  o << setw(15) << d.Pav.GetMeanError() / d.rho / T0 << " " << "\n";
}

constexpr long kChunk = 1000;  // steps per copy of the per-step history

}  // namespace

int main(int argc, char* argv[])
{
  Parameters params;
  params.output_prefix = "isotherm.run";
  params.parameters = {{"N", 400},        {"T*", 1.4},   {"rho*_min", 0.60}, {"rho*_max", 0.60},
                       {"drho*", 0.01},   {"teq", 10.},  {"tfin", 5000.},    {"dt*", 0.004},
                       {"useCUDA", 1}};
  if (argc > 1) {
    cout << "Reading parameters from file " << std::string(argv[1]) << endl;
    params.ReadParametersFromFile(std::string(argv[1]));
  }
  {
    // GetFullPrefix (run-isotherm-aux.h:99-117)
    std::stringstream ss;
    const std::string ts = params.timestamp();
    if (!ts.empty()) ss << "." << ts;
    ss << ".N" << params.parameters["N"] << ".Tst" << params.parameters["T*"];
    params.output_prefix += ss.str();
  }
  require_device(params);
  params.OutputParameters("isotherm");

  const int N = (int)params.iget("N");
  const double Tst = params.get("T*");
  const double rhomin = params.get("rho*_min"), rhomax = params.get("rho*_max");
  const double drho = params.get("drho*");
  const double dt = params.get("dt*"), teq = params.get("teq"), tfin = params.get("tfin");
  const uint64_t seed = params.has("seed") ? (uint64_t)params.iget("seed") : 1;
  const bool batched = !params.has("batched") || params.iget("batched") != 0;
  const long every = params.has("print_every") ? std::max(1L, params.iget("print_every")) : 1000;
  if (N <= 0 || !(dt > 0.) || !(drho > 0.)) die("N, dt* and drho* must be positive");

  std::vector<std::unique_ptr<Density>> dens;
  for (double rho = rhomin; rho <= rhomax; rho += drho) {  // :84
    dens.emplace_back(new Density());
    dens.back()->rho = rho;
  }
  const int nb = (int)dens.size();
  double t = 0.;
  const long neq = count_steps(t, teq, dt);   // :100-103
  const double t_prod = t;
  const long nprod = count_steps(t, tfin, dt);  // :109-133
  const double t_end = t;

  cout << setw(15) << "rho*" << " " << setw(15) << "t*" << " " << setw(15) << "T*" << " ";
  cout << setw(15) << "<u*>" << " " << setw(15) << "d<u*>" << " " << setw(15) << "teq_u*" << " ";
  cout << setw(15) << "<p*>" << " " << setw(15) << "d<p*>" << " " << setw(15) << "teq_p*" << " ";
  cout << setw(15) << "<Z>" << " " << setw(15) << "d<Z>" << " " << endl;

  std::ofstream fout(params.output_prefix + ".dat");
  fout << setw(15) << "rho*" << " " << setw(15) << "t*" << " " << setw(15) << "T*" << " ";
  fout << setw(15) << "<u*>" << " " << setw(15) << "d<u*>" << " " << setw(15) << "teq_u*" << " ";
  fout << setw(15) << "<p*>" << " " << setw(15) << "d<p*>" << " " << setw(15) << "teq_p*" << " ";
  fout << setw(15) << "<Z>" << " " << setw(15) << "d<Z>" << " " << setw(15) << "w[N]" << " ";
  fout << setw(15) << "dw[N]" << " " << endl;

  if (batched && nb > 0) {
    std::vector<int> Ns(nb, N);
    std::vector<double> rhos(nb), T0s(nb, Tst);
    for (int k = 0; k < nb; ++k) rhos[k] = dens[k]->rho;
    lj_batch* b = nullptr;
    check(lj_batch_create(nb, Ns.data(), rhos.data(), T0s.data(), LJ_TVN, LJ_BC_PERIODIC, 0, &b),
          "lj_batch_create");
    check(lj_batch_sample_initial_state(b, seed), "lj_batch_sample_initial_state");
    for (long done = 0; done < neq;) {
      const long n = std::min(kChunk * 10, neq - done);
      check(lj_batch_integrate(b, dt, (int)n), "lj_batch_integrate");
      done += n;
    }
    std::vector<lj_obs> obs(nb);
    check(lj_batch_get_observables(b, obs.data()), "lj_batch_get_observables");
    std::vector<double> hist((size_t)kChunk * nb * 5);
    double tt = t_prod;
    long iter = 0;
    // the observation before step i is the state after step i - 1
    std::vector<double> U(nb), P(nb), T(nb);
    for (int k = 0; k < nb; ++k) {
      U[k] = obs[k].U;
      P[k] = obs[k].P;
      T[k] = obs[k].T;
    }
    for (long done = 0; done < nprod;) {
      const long n = std::min(kChunk, nprod - done);
      check(lj_batch_integrate_hist(b, dt, (int)n, hist.data()), "lj_batch_integrate_hist");
      for (long s = 0; s < n; ++s) {
        ++iter;
        for (int k = 0; k < nb; ++k)
          observe(*dens[k], N, U[k], P[k], T[k], tt, iter, every, dt, Tst);
        const double* h = hist.data() + (size_t)s * nb * 5;
        for (int k = 0; k < nb; ++k) {
          U[k] = h[5 * k + 0];
          T[k] = h[5 * k + 1];
          P[k] = h[5 * k + 4];
        }
        tt += dt;
      }
      done += n;
    }
    for (int k = 0; k < nb; ++k) dens[k]->T_end = nprod > 0 ? T[k] : obs[k].T;
    lj_batch_destroy(b);
  } else {
    for (int k = 0; k < nb; ++k) {
      Density& d = *dens[k];
      lj_config c;
      lj_config_init(&c);
      c.N = N;
      c.T0 = Tst;
      c.rho = d.rho;
      c.canonical = LJ_TVN;  // :92
      c.boundary = LJ_BC_PERIODIC;
      lj_ctx* ctx = nullptr;
      check(lj_create(&c, &ctx), "lj_create");
      check(lj_sample_initial_state(ctx, seed + (uint64_t)k), "lj_sample_initial_state");
      for (long done = 0; done < neq;) {
        const long n = std::min(kChunk * 10, neq - done);
        check(lj_integrate(ctx, dt, (int)n, nullptr), "lj_integrate");
        done += n;
      }
      lj_obs o;
      check(lj_get_observables(ctx, &o), "lj_get_observables");
      double U = o.U, P = o.P, T = o.T, tt = t_prod;
      std::vector<double> hist((size_t)kChunk * 5);
      long iter = 0;
      for (long done = 0; done < nprod;) {
        const long n = std::min(kChunk, nprod - done);
        check(lj_integrate(ctx, dt, (int)n, hist.data()), "lj_integrate");
        for (long s = 0; s < n; ++s) {
          ++iter;
          observe(d, N, U, P, T, tt, iter, every, dt, Tst);
          U = hist[5 * s + 0];
          T = hist[5 * s + 1];
          P = hist[5 * s + 4];
          tt += dt;
        }
        done += n;
      }
      d.T_end = T;
      lj_destroy(ctx);
    }
  }

  // console rows and the results file in density order (:117-165)
  double Pprev = 0., Ppreverr = 0.;
  for (int k = 0; k < nb; ++k) {
    Density& d = *dens[k];
    cout << endl << d.rows.str();
    cout.flush();
    fout << setw(15) << d.rho << " " << setw(15) << t_end - dt << " " << setw(15) << d.T_end << " ";
    fout << setw(15) << d.uav.GetMean() << " " << setw(15) << d.uav.GetMeanError() << " ";
    fout << setw(15) << d.uav.GetS() * dt << " ";
    fout << setw(15) << d.Pav.GetMean() << " " << setw(15) << d.Pav.GetMeanError() << " ";
    fout << setw(15) << d.Pav.GetS() * dt << " ";
    fout << setw(15) << d.Pav.GetMean() / d.rho / Tst << " ";
    fout << setw(15) << d.Pav.GetMeanError() / d.rho / Tst << " ";
    const double Pcur = d.Pav.GetMean(), Pcurerr = d.Pav.GetMeanError();
    double tdrho = drho;
    if (d.rho == rhomin) tdrho = rhomin;
    const double w = Tst / ((Pcur - Pprev) / tdrho);
    const double werr =
        Tst * tdrho * std::sqrt(Pcurerr * Pcurerr + Ppreverr * Ppreverr) / std::fabs(Pcur - Pprev);
    fout << setw(15) << w << " " << setw(15) << werr << " ";
    Pprev = Pcur;
    Ppreverr = Pcurerr;
    fout << endl;
    fout.flush();
  }
  return 0;
}
