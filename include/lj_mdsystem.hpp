/* lj_mdsystem.hpp -- the reference's simulation class (MDSystem.h:6-164) as a
 * header-only C++ class over the C-ABI of libljhip.so (lj_hip.h, group 2).
 *
 * Same member names, public fields and method semantics as the reference
 * class, with the state kept on the GPU: Integrate() replays a captured HIP
 * graph of the whole step instead of the reference's per-step host
 * integrator and PCIe round trip (MDSystem.cpp:240-250,438-583).
 *
 * Differences from the reference class, each deliberate:
 *   - the state lives on the device.  With host sync on (the default) every
 *     Integrate/CalculateForces call refreshes h_Pos, h_Vel, h_Force and the
 *     observables, as the reference's GPU branch leaves them; setHostSync(false)
 *     skips the array copies (fields are still refreshed);
 *   - initial velocities come from a seeded generator (m_config.seed), not
 *     from the clock-seeded MTRand (MersenneTwister.h:262): by default the
 *     benchmark configurations' Gaussian draw, or with
 *     m_config.initialVelocities = LJ_INIT_MAXWELL the reference's own
 *     Maxwell-speed rejection sampler with isotropic directions;
 *   - setArray writes the device state, which here is authoritative (on the
 *     reference's GPU branch the next CalculateForces overwrites it from h_Pos);
 *   - NdNdr2 is refreshed by CalculateForces() and, as in the reference,
 *     by every step of Integrate(): each step bins the positions its force
 *     pass sees (an O(N^2) pass per step, on the device).  setRDFEveryStep(false)
 *     drops that pass; NdNdr2 then changes only on CalculateForces() and
 *     UpdateRDF().  RDF() and getvelo() return their curves as two vectors
 *     instead of a SplineFunction;
 *   - there is no CPU branch: setHardwareMode(false) throws.  The reference's
 *     own class linked against libljhip.so keeps its CPU/GPU switch intact;
 *   - at N >= 3072 (FP32) h_Force[4 i + 3] is 0 after Integrate (the step sums
 *     the potential energy per workgroup; V, P are unaffected) and is
 *     refreshed by CalculateForces();
 *   - Pshear stays 0, as on the reference's GPU branch; UpdateShearPressure()
 *     sets it to the CPU path's value for the current state;
 *   - every failure throws std::runtime_error with lj_last_error().
 *   Fluctuations (MDSystem.cpp:708-730) counts the sub-volume on the device
 *   (lj_fl_*) and keeps the reference's moment sums and printed line.
 */
#pragma once

#include <cmath>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lj_hip.h"

namespace ljhip {

class MDSystem {
 public:
  struct MDSystemConfiguration {  // MDSystem.h:10-41, defaults as there
    int N;
    double T0;
    double rho;
    bool canonical;
    int boundaryConditions;  // 0 periodic, 1 hard wall, 2 none
    bool useCUDA;            // the GPU switch; only true is supported here
    int CUDABlockSize;       // accepted, unused (the kernels pick their own)
    double eps, sigma;
    int precision;           // LJ_PREC_*
    int device;
    uint64_t seed;
    // LJ_INIT_GAUSSIAN (the benchmark configurations' initial velocities) or
    // LJ_INIT_MAXWELL (the reference class's own sampler, MDSystem.cpp:147-181)
    int initialVelocities;
    MDSystemConfiguration()
        : N(128), T0(1.5), rho(0.2), canonical(false), boundaryConditions(0), useCUDA(true),
          CUDABlockSize(256), eps(1.0), sigma(1.0), precision(LJ_PREC_FP32), device(0), seed(1),
          initialVelocities(LJ_INIT_GAUSSIAN)
    {
    }
  };

  struct Curve {  // the points of the reference's SplineFunction results
    std::vector<double> x, y;
  };

  MDSystemConfiguration m_config;
  double U = 0, T = 0, K = 0, V = 0, P = 0;
  double Pshear = 0;
  float* h_Pos = nullptr;
  float* h_Vel = nullptr;
  float* h_Force = nullptr;
  float* d_Pos = nullptr;  // device arrays of the context (AoS float4)
  float* d_Vel = nullptr;
  float* d_Force = nullptr;
  double t = 0;
  double L = 0;
  std::vector<int> NdNdr2;
  float rdf_dr2 = 0.f;
  std::vector<std::pair<double, double> > curvelo;  // speed distribution (MDSystem.h:86)
  int veloIters = 0;
  double av_U_tot = 0, av_T_tot = 0, av_p_tot = 0;
  int av_iters = 0;
  double momN = 0, momN2 = 0, momN3 = 0, momN4 = 0;  // MDSystem.h:94-95 (Fluctuations)
  int momiters = 0;

  explicit MDSystem(const MDSystemConfiguration& config = MDSystemConfiguration())
  {
    Reinitialize(config);
  }
  ~MDSystem() { release(); }
  MDSystem(const MDSystem&) = delete;
  MDSystem& operator=(const MDSystem&) = delete;

  // MDSystem.cpp:66-114
  void Reinitialize(const MDSystemConfiguration& config)
  {
    if (!config.useCUDA) throw std::runtime_error(no_cpu());
    release();
    m_config = config;
    lj_config c;
    lj_config_init(&c);
    c.N = config.N;
    c.T0 = config.T0;
    c.rho = config.rho;
    c.canonical = config.canonical ? LJ_TVN : LJ_EVN;
    c.boundary = config.boundaryConditions;
    c.eps = config.eps;
    c.sigma = config.sigma;
    c.precision = config.precision;
    c.device = config.device;
    check(lj_create(&c, &ctx_));
    L = lj_box_length(ctx_);
    const size_t n4 = (size_t)4 * config.N;
    h_Pos = new float[n4];
    h_Vel = new float[n4];
    h_Force = new float[n4];
    check(lj_device_arrays(ctx_, &d_Pos, &d_Vel, &d_Force));
    rdf_dr2 = lj_rdf_dr2_default(config.N);
    NdNdr2.assign(256, 0);
    momN = momN2 = momN3 = momN4 = 0.;  // MDSystem.cpp:109-110
    momiters = 0;
    fl_fraction_ = -1.;
    // the step's RDF pass runs in the FP32 modes
    step_rdf_on_ = false;
    if (rdf_every_step_ && config.precision != LJ_PREC_FP64) enable_step_rdf(true);
    seed_ = config.seed;
    init_velocities_ = config.initialVelocities;
    SampleInitialConditions();
    CalculateForces();
    CalculateParameters();
    veloIters = 0;
    initvelo();
    resetAveraging();
  }

  // MDSystem.cpp:147-181 (lattice, then seeded Gaussian or Maxwell velocities,
  // m_config.initialVelocities, with the momentum removed and T rescaled to T0)
  void SampleInitialConditions()
  {
    check(lj_sample_initial_state_ex(ctx_, seed_++, init_velocities_));
    t = 0;
    pull_fields();
    pull_arrays();
  }

  // MDSystem.cpp:183-216 on the device velocities (host round trip)
  void CorrectTotalMomentum()
  {
    getArray(1);
    double px = 0, py = 0, pz = 0, m = 0;
    for (int i = 0; i < 4 * m_config.N; i += 4) {
      px += h_Vel[i];
      py += h_Vel[i + 1];
      pz += h_Vel[i + 2];
      m += 1.;
    }
    for (int i = 0; i < 4 * m_config.N; i += 4) {
      h_Vel[i] += -px / m;
      h_Vel[i + 1] += -py / m;
      h_Vel[i + 2] += -pz / m;
    }
    setArray(1, h_Vel);
  }

  // MDSystem.cpp:232-311: forces, V and the virial part of P; NdNdr2 with the
  // reference GPU binning (MDSystem.cu:39)
  void CalculateForces()
  {
    lj_obs o;
    check(lj_compute_forces(ctx_, &o));
    V = o.V;
    P = o.W / 3.;
    UpdateRDF();
    if (sync_) pull_arrays();
  }

  // MDSystem.cpp:325-359
  void CalculateParameters()
  {
    lj_obs o;
    check(lj_compute_parameters(ctx_, &o));
    set_fields(o);
  }

  double KineticTemperature()  // MDSystem.cpp:361-373 on the current velocities
  {
    lj_obs o;
    check(lj_compute_parameters(ctx_, &o));
    return o.T;
  }

  void RenormalizeVelocities(bool RecalculateTkin = false)  // MDSystem.cpp:375-389
  {
    check(lj_renormalize_velocities(ctx_, RecalculateTkin ? 1 : 0));
    pull_fields();
    if (sync_) pull_arrays();
  }

  void RenormalizeVelocitiesToEnergy(double ust)  // MDSystem.cpp:391-404
  {
    check(lj_renormalize_to_energy(ctx_, ust));
    pull_fields();
    if (sync_) pull_arrays();
  }

  void ApplyBoundaryConditions()  // MDSystem.cpp:406-436
  {
    check(lj_apply_boundary_conditions(ctx_));
    if (sync_) pull_arrays();
  }

  double getTime() { return t; }

  // MDSystem.cpp:438-583: one step
  void Integrate(double dt) { Integrate(dt, 1); }

  // nsteps steps back to back on the device; fields and arrays refreshed once
  void Integrate(double dt, int nsteps)
  {
    check(lj_integrate(ctx_, dt, nsteps, nullptr));
    pull_fields();
    if (step_rdf_on_ && nsteps > 0) pull_step_rdf();
    if (sync_) pull_arrays();
  }

  // the RDF pass inside every step (default on, as the reference fills
  // NdNdr2 in every CalculateForces, MDSystem.cpp:254,281-285)
  void setRDFEveryStep(bool on)
  {
    rdf_every_step_ = on;
    enable_step_rdf(on && m_config.precision != LJ_PREC_FP64);
  }

  // MDSystem.cpp:708-730: particles in the closed centred cube of volume
  // fraction `fraction` of the current positions, added to the moment sums,
  // and the reference's progress line on stdout
  void Fluctuations(double fraction)
  {
    if (fraction != fl_fraction_) {
      lj_fl_ladder l;
      l.type = LJ_FL_CUBE_CLOSED;
      l.single = 1;
      l.alpha = fraction;
      l.vmax = 0.;
      check(lj_fl_configure(ctx_, &l, 1));
      fl_fraction_ = fraction;
    }
    check(lj_fl_sample(ctx_));
    int64_t c = 0;
    check(lj_fl_counts(ctx_, &c));
    const long long tN = (long long)c;
    const int N = m_config.N;
    momN += tN;
    momN2 += tN * tN;
    momN3 += tN * tN * tN;
    momN4 += tN * tN * tN * tN;
    momiters++;
    std::cout << "Iteration: " << momiters << "\t" << "<N> = " << momN / momiters << "\t"
              << "w[N] = " << (momN2 / momiters - (momN / momiters) * (momN / momiters)) / (momN / momiters)
              << "\t" << "EV: " << (1. - 2. * pi() / 3. * N / L / L / L) * (1. - 2. * pi() / 3. * N / L / L / L)
              << "\t" << "Binom: " << 1. - fraction << "\n";
  }

  // MDSystem.cpp:586-630 (type 0 positions, 1 velocities; 2 forces added)
  float* getArray(int type)
  {
    float* h = type == 1 ? h_Vel : type == 2 ? h_Force : h_Pos;
    check(lj_get_array(ctx_, type == 1 ? LJ_ARRAY_VEL : type == 2 ? LJ_ARRAY_FORCE : LJ_ARRAY_POS, h));
    return h;
  }

  void setArray(int type, const float* data)
  {
    check(lj_set_array(ctx_, type == 1 ? LJ_ARRAY_VEL : type == 2 ? LJ_ARRAY_FORCE : LJ_ARRAY_POS,
                       data));
  }

  void resetAveraging()  // MDSystem.cpp:701-705
  {
    check(lj_reset_averaging(ctx_));
    av_U_tot = av_T_tot = av_p_tot = 0.;
    av_iters = 0;
  }

  // Pshear as the reference's CPU path leaves it after CalculateForces and
  // CalculateParameters (MDSystem.cpp:299,309,335,353), for the current state.
  // The GPU branch, which this class follows, never computes it (Pshear stays
  // 0 after Integrate and CalculateParameters); this runs one all-pairs pass.
  double UpdateShearPressure()
  {
    check(lj_shear_pressure(ctx_, &Pshear));
    return Pshear;
  }

  // NdNdr2 of the current positions (256 r^2 bins of width rdf_dr2)
  void UpdateRDF()
  {
    std::vector<uint64_t> c(NdNdr2.size());
    check(lj_rdf(ctx_, LJ_RDF_REF_GPU, (int)c.size(), rdf_dr2, 0., c.data(), 0));
    for (size_t k = 0; k < c.size(); ++k) NdNdr2[k] = (int)c[k];
  }

  // MDSystem.cpp:633-649: the normalised g(r) of NdNdr2 (arguments unused, as there)
  Curve RDF(double /*rmax*/ = 5., double /*shag*/ = 0.1) const
  {
    const int N = m_config.N;
    const double n0 = N / L / L / L;
    Curve out;
    for (size_t ir = 0; ir < NdNdr2.size(); ++ir) {
      const double r = std::sqrt((ir + 0.5) * rdf_dr2);
      out.x.push_back(r);
      out.y.push_back(NdNdr2[ir] / rdf_dr2 / 2. / pi() / r / n0 / static_cast<double>(N));
    }
    return out;
  }

  // MDSystem.cpp:651-669: speed distribution of the current velocities,
  // maxind = (int)(vmax / step) + 1 points at step (i + 1/2)
  void initvelo(double vmax = 12., double step = 0.12)
  {
    const int maxind = (int)(vmax / step) + 1;
    std::vector<uint64_t> dens = speed_counts(step, maxind);
    curvelo.clear();
    for (int i = 0; i < maxind; ++i)
      curvelo.push_back(std::make_pair(step * (0.5 + i), (int)dens[i] / step / m_config.N));
    veloIters = 1;
  }

  // MDSystem.cpp:676-694: running average of the distribution
  void updatevelo()
  {
    if (curvelo.size() < 2) throw std::runtime_error("ljhip: updatevelo before initvelo");
    const int maxind = (int)curvelo.size();
    const double shag = curvelo[1].first - curvelo[0].first;
    std::vector<uint64_t> dens = speed_counts(shag, maxind);
    for (int i = 0; i < maxind; ++i)
      curvelo[i].second =
          (curvelo[i].second * veloIters + (int)dens[i] / shag / m_config.N) / (veloIters + 1);
    veloIters++;
  }

  Curve getvelo() const  // MDSystem.cpp:671-674
  {
    Curve out;
    for (size_t i = 0; i < curvelo.size(); ++i) {
      out.x.push_back(curvelo[i].first);
      out.y.push_back(curvelo[i].second);
    }
    return out;
  }

  double Maxwell(double v)  // MDSystem.cpp:696-699, at the current temperature T
  {
    return 4 * pi() * std::pow(1. / 2. / pi() / T, 3. / 2.) * v * v * std::exp(-v * v / 2. / T);
  }

  int fast_round(float x)  // MDSystem.cpp:732-739
  {
    return x > 0 ? static_cast<int>(x + 0.5f) : static_cast<int>(x - 0.5f);
  }

  void setHardwareMode(bool useCUDA)
  {
    if (!useCUDA) throw std::runtime_error(no_cpu());
  }
  void setCanonical(bool canonical)
  {
    m_config.canonical = canonical;
    apply_mode();
  }
  void setPeriodicBoundaryCondition(bool periodic)
  {
    m_config.boundaryConditions = !periodic;
    apply_mode();
  }
  void setBoundaryCondition(int cond)
  {
    m_config.boundaryConditions = cond;
    apply_mode();
  }

  // host mirrors refreshed after each call (default true)
  void setHostSync(bool on) { sync_ = on; }
  lj_ctx* context() { return ctx_; }

 private:
  static const char* no_cpu()
  {
    return "ljhip::MDSystem runs on the GPU only; the reference class keeps the CPU branch";
  }
  lj_ctx* ctx_ = nullptr;
  bool sync_ = true;
  bool rdf_every_step_ = true;
  bool step_rdf_on_ = false;
  double fl_fraction_ = -1.;  // fraction the device counter is set up for
  uint64_t seed_ = 1;
  int init_velocities_ = LJ_INIT_GAUSSIAN;

  void enable_step_rdf(bool on)
  {
    lj_step_sampling smp;
    smp.rdf = on ? LJ_SAMPLE_RDF_LAST : LJ_SAMPLE_RDF_OFF;
    smp.rdf_rule = LJ_RDF_REF_GPU;
    smp.rdf_nbins = (int)NdNdr2.size();
    smp.rdf_dr2 = rdf_dr2;
    smp.rdf_rmax = 0.;
    smp.fluctuations = 0;
    check(lj_set_step_sampling(ctx_, &smp));
    step_rdf_on_ = on;
  }

  void pull_step_rdf()
  {
    std::vector<uint64_t> c(NdNdr2.size());
    check(lj_step_rdf_counts(ctx_, c.data(), nullptr, 0));
    for (size_t k = 0; k < c.size(); ++k) NdNdr2[k] = (int)c[k];
  }

  static double pi() { return 3.141592653589793238462643; }  // MDSystem.cpp:7

  std::vector<uint64_t> speed_counts(double step, int nbins)
  {
    if (nbins > 16384) throw std::runtime_error("ljhip: speed histogram limited to 16384 bins");
    std::vector<uint64_t> c((size_t)nbins);
    check(lj_velocity_histogram(ctx_, step, nbins, c.data()));
    return c;
  }

  static void check(int rc)
  {
    if (rc != LJ_OK) throw std::runtime_error(std::string("ljhip: ") + lj_last_error());
  }

  void release()
  {
    if (ctx_) lj_destroy(ctx_);
    ctx_ = nullptr;
    delete[] h_Pos;
    delete[] h_Vel;
    delete[] h_Force;
    h_Pos = h_Vel = h_Force = nullptr;
    d_Pos = d_Vel = d_Force = nullptr;
  }

  void apply_mode()
  {
    check(lj_set_mode(ctx_, m_config.canonical ? LJ_TVN : LJ_EVN, m_config.boundaryConditions));
  }

  void set_fields(const lj_obs& o)
  {
    U = o.U;
    T = o.T;
    K = o.K;
    V = o.V;
    P = o.P;
    Pshear = o.Pshear;
    t = o.t;
    av_U_tot = o.av_U_tot;
    av_T_tot = o.av_T_tot;
    av_p_tot = o.av_p_tot;
    av_iters = (int)o.av_iters;
  }

  void pull_fields()
  {
    lj_obs o;
    check(lj_get_observables(ctx_, &o));
    set_fields(o);
  }

  void pull_arrays()
  {
    check(lj_get_array(ctx_, LJ_ARRAY_POS, h_Pos));
    check(lj_get_array(ctx_, LJ_ARRAY_VEL, h_Vel));
    check(lj_get_array(ctx_, LJ_ARRAY_FORCE, h_Force));
  }
};

}  // namespace ljhip
