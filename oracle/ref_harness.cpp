// ref_harness.cpp -- C entry points over the reference's own MDSystem class.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference sources
// where they lie (src/library/MDSystem.cpp, splinefunction.cpp of
// vlvovch/lennard-jones-cuda) by oracle/Makefile into oracle/_ref/:
//   libref_cpu.so  -- reference built without USE_CUDA_TOOLKIT: its CPU path
//                     (MDSystem.cpp:252-310) is the test reference and CPU baseline;
//   libref_hip.so  -- reference built with USE_CUDA_TOOLKIT and linked
//                     against this project's libljhip.so, so that the
//                     reference's own CPU/GPU switch (MDSystem.cpp:240) drives
//                     the HIP path through the compatibility C-ABI.
// The reference initialises from a clock-seeded RNG (MersenneTwister.h:262),
// so every harness user injects its own state through h_Pos / h_Vel
// (public members, MDSystem.h:72-74) before computing anything.
#include <cstring>
#include "MDSystem.h"

extern "C" {

void* refh_create(int N, double T0, double rho, int canonical, int bc, int useCUDA,
                  int blockSize)
{
  MDSystem::MDSystemConfiguration cfg;  // MDSystem.h:10-41
  cfg.N = N;
  cfg.T0 = T0;
  cfg.rho = rho;
  cfg.canonical = canonical != 0;
  cfg.boundaryConditions = bc;
  cfg.useCUDA = useCUDA != 0;
  if (blockSize > 0) cfg.CUDABlockSize = blockSize;
  return new MDSystem(cfg);  // MDSystem.cpp:55-64 (runs one CalculateForces)
}

void refh_destroy(void* h) { delete static_cast<MDSystem*>(h); }

void refh_set_state(void* h, const float* pos4, const float* vel4)
{
  MDSystem* s = static_cast<MDSystem*>(h);
  const int n4 = 4 * s->m_config.N;
  if (pos4) std::memcpy(s->h_Pos, pos4, sizeof(float) * n4);
  if (vel4) std::memcpy(s->h_Vel, vel4, sizeof(float) * n4);
  s->t = 0.;
  s->resetAveraging();
}

void refh_set_force(void* h, const float* frc4)
{
  MDSystem* s = static_cast<MDSystem*>(h);
  std::memcpy(s->h_Force, frc4, sizeof(float) * 4 * s->m_config.N);
}

// CalculateForces() (MDSystem.cpp:232-311); out3 = {V, P, Pshear}
void refh_forces(void* h, float* frc4, double* out3, int* rdf256)
{
  MDSystem* s = static_cast<MDSystem*>(h);
  s->CalculateForces();
  if (frc4) std::memcpy(frc4, s->h_Force, sizeof(float) * 4 * s->m_config.N);
  if (out3) {
    out3[0] = s->V;
    out3[1] = s->P;
    out3[2] = s->Pshear;
  }
  if (rdf256)
    for (int b = 0; b < 256; ++b) rdf256[b] = s->NdNdr2[b];
}

// CalculateParameters() (MDSystem.cpp:325-359)
void refh_params(void* h) { static_cast<MDSystem*>(h)->CalculateParameters(); }

// out8 = {U, T, K, V, P, Pshear, t, L}
void refh_observables(void* h, double* out8)
{
  MDSystem* s = static_cast<MDSystem*>(h);
  out8[0] = s->U;
  out8[1] = s->T;
  out8[2] = s->K;
  out8[3] = s->V;
  out8[4] = s->P;
  out8[5] = s->Pshear;
  out8[6] = s->t;
  out8[7] = s->L;
}

// n calls of Integrate(dt) (MDSystem.cpp:438-583); per-step {U,T,K,V,P} into
// hist (5 doubles per step) when hist != NULL
void refh_integrate(void* h, double dt, int nsteps, double* hist)
{
  MDSystem* s = static_cast<MDSystem*>(h);
  for (int k = 0; k < nsteps; ++k) {
    s->Integrate(dt);
    if (hist) {
      hist[5 * k + 0] = s->U;
      hist[5 * k + 1] = s->T;
      hist[5 * k + 2] = s->K;
      hist[5 * k + 3] = s->V;
      hist[5 * k + 4] = s->P;
    }
  }
}

void refh_get_state(void* h, float* pos4, float* vel4, float* frc4)
{
  MDSystem* s = static_cast<MDSystem*>(h);
  const int n4 = 4 * s->m_config.N;
  if (pos4) std::memcpy(pos4, s->h_Pos, sizeof(float) * n4);
  if (vel4) std::memcpy(vel4, s->h_Vel, sizeof(float) * n4);
  if (frc4) std::memcpy(frc4, s->h_Force, sizeof(float) * n4);
}

double refh_L(void* h) { return static_cast<MDSystem*>(h)->L; }
float refh_dr2(void* h) { return static_cast<MDSystem*>(h)->rdf_dr2; }
int refh_fast_round(void* h, float x) { return static_cast<MDSystem*>(h)->fast_round(x); }

// RDF() normalisation (MDSystem.cpp:633-649); x, y of 256 points
void refh_rdf_normalised(void* h, double* x, double* y)
{
  SplineFunction f = static_cast<MDSystem*>(h)->RDF();
  for (size_t k = 0; k < f.vals.size() && k < 256; ++k) {
    x[k] = f.vals[k].first;
    y[k] = f.vals[k].second;
  }
}

// velocity distribution (MDSystem.cpp:651-694): initvelo on the current
// velocities, or updatevelo (running average) when update != 0; returns the
// number of curve points written to x, y (at most cap)
int refh_velo(void* h, int update, double vmax, double step, double* x, double* y, int cap)
{
  MDSystem* m = static_cast<MDSystem*>(h);
  if (update)
    m->updatevelo();
  else
    m->initvelo(vmax, step);
  const int n = (int)m->curvelo.vals.size();
  for (int k = 0; k < n && k < cap; ++k) {
    x[k] = m->curvelo.vals[k].first;
    y[k] = m->curvelo.vals[k].second;
  }
  return n;
}

double refh_maxwell(void* h, double v) { return static_cast<MDSystem*>(h)->Maxwell(v); }

}  // extern "C"
