/*
 * lj_oracle.h -- CPU restatement of the reference Lennard-Jones time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the test oracle of the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only to produce expected values or a timed CPU baseline.  The
 * product library (lennard-jones-cuda_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates
 * (vlvovch/lennard-jones-cuda, src/library/MDSystem.cpp unless noted).
 * Arithmetic types and evaluation order follow the reference exactly
 * (float state, double pair math, float force accumulation), so that with
 * the same inputs the restatement is bit-identical to the reference CPU path;
 * build with -ffp-contract=off and without -ffast-math.
 *
 * Parity pins (see tests/test_oracle_*.py): bit-exact against the reference
 * CPU path compiled from source into oracle/_ref (when present); closed forms
 * (LJ dimer, periodic simple-cubic lattice), F = -grad V by central
 * differences, Newton's third law, virial identity, brute-force integer
 * RDF counts, time reversibility and the TVN half-step invariant.
 * The seeded Gaussian initialiser (orc_gaussian_velocities) has no reference
 * counterpart (the reference seeds from the clock); it is pinned by its
 * invariants (zero momentum, exact kinetic temperature) and by the closed
 * form of SplitMix64 / Box-Muller on the first draws.
 */
#ifndef LJ_ORACLE_H
#define LJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- initial state ---------------- */

/* lattice sites per edge: ceil(pow(N, 1/3))            MDSystem.cpp:150 */
int    orc_lattice_n(int N);
/* box side L = (N/rho)^(1/3)                           MDSystem.cpp:70 */
double orc_box_length(int N, double rho);
/* RDF bin width in r^2 (reference rule, float)         MDSystem.cpp:93-95 */
float  orc_rdf_dr2(int N);
/* simple-cubic lattice, particle k on site k, x fastest MDSystem.cpp:150-168 */
void   orc_lattice_positions(int N, double L, float* pos4);
/* seeded Gaussian v_c = sqrt(T0) g, g ~ N(0,1) from SplitMix64 + Box-Muller,
 * then momentum removal and exact rescale to T0 (reference steps :179-180). */
void   orc_gaussian_velocities(int N, double T0, uint64_t seed, float* vel4);
/* uniform double in [0,1) from SplitMix64 state (advances *state) */
double orc_splitmix_uniform(uint64_t* state);
/* total momentum removal                                MDSystem.cpp:183-216 */
void   orc_correct_momentum(int N, float* vel4);
/* T_kin = sum v^2 / (3N), float products, double sum   MDSystem.cpp:361-373 */
double orc_kinetic_temperature(int N, const float* vel4);
/* v *= sqrt(T0/Tkin)                                    MDSystem.cpp:375-389 */
void   orc_renormalize_velocities(int N, double T0, double Tkin, float* vel4);
/* (int)(x +- 0.5f)                                      MDSystem.cpp:732-739 */
int    orc_fast_round(float x);

/* ---------------- forces ---------------- */

/* Reference CPU force/energy/virial/RDF pass, bit-faithful.
 * MDSystem.cpp:237-311 (CPU branch).  frc4[4i+3] is left untouched, as in
 * the reference.  rdf256 receives 256 counts (r^2 bins of width dr2). */
void orc_forces_ref(int N, const float* pos4, double L, int bc, float dr2,
                    float* frc4, double* V, double* P, double* Pshear,
                    int* rdf256);

/* Same as orc_forces_ref but parallel over i with nthreads OpenMP threads.
 * Forces and RDF are bit-identical to orc_forces_ref (per-i j order is
 * unchanged); V, P, Pshear are summed per i, then over i in order. */
void orc_forces_ref_omp(int N, const float* pos4, double L, int bc, float dr2,
                        float* frc4, double* V, double* P, double* Pshear,
                        int* rdf256, int nthreads);

/* FP64 all-pairs oracle with LJ parameters eps, sigma.
 *   f3[3i..]   = sum_j 24 eps (2 (s/r)^12 - (s/r)^6) r_ij / r^2     (force)
 *   ew[2i]     = sum_j ((s/r)^12 - (s/r)^6)         (the Force.w convention,
 *                 MDSystem.cu:52, unscaled)
 *   ew[2i+1]   = sum_j r_ij . f_ij                   (per-i virial)
 *   *E = 4 eps sum_{i<j} ((s/r)^12-(s/r)^6),  *W = sum_{i<j} r_ij . f_ij.
 * periodic != 0 -> minimum image r -= L round(r/L) (half away from zero, the
 * reference's tie rule MDSystem.cpp:732-739).
 * skip_r2 <= 0: exclude the self pair by index (MDSystem.cpp:267);
 * skip_r2 > 0: exclude every pair with r^2 <= skip_r2 (MDSystem.cu:37).
 * Any output pointer may be NULL.  nthreads <= 0 uses all cores. */
void orc_forces_fp64(int N, const float* pos4, double L, int periodic,
                     double eps, double sigma, double skip_r2,
                     double* f3, double* ew, double* E, double* W,
                     int nthreads);
/* the same from double positions (3 per particle) */
void orc_forces_fp64d(int N, const double* pos3, double L, int periodic,
                      double eps, double sigma,
                      double* f3, double* ew, double* E, double* W,
                      int nthreads);

/* ---------------- RDF ---------------- */

enum {
  ORC_RDF_REF_CPU = 0, /* MDSystem.cpp:273-285: double-faithful min image,
                          exact floor(r2/dr2), drop bins >= nbins, self by index */
  ORC_RDF_REF_GPU = 1, /* MDSystem.cu:26-42: float lround min image,
                          r2 = fma(z,z,fma(y,y,x*x)), floor(fl32(r2/dr2)),
                          clamp to nbins-1, skip r2 <= 1e-6f */
  ORC_RDF_RBINS   = 2  /* r^2 as REF_CPU; nbins uniform in r on [0, rmax):
                          bin = max k with edges2[k] <= r2, edges2[k] =
                          (float)((k*(rmax/nbins))^2); drop r2 >= edges2[nbins] */
};
/* counts[nbins] accumulated (not cleared) over ordered pairs. */
void orc_rdf(int N, const float* pos4, double L, int periodic, int rule,
             int nbins, float dr2, double rmax, int64_t* counts, int nthreads);
/* squared bin edges for ORC_RDF_RBINS, edges2[0..nbins] */
void orc_rdf_edges2(int nbins, double rmax, float* edges2);

/* ---------------- full system (integrator, BCs, observables) ---------------- */

typedef struct {
  int N;
  double T0, rho;
  int canonical;   /* 0 EVN, 1 TVN                      MDSystem.h:21 */
  int bc;          /* 0 periodic, 1 hard wall, 2 none   MDSystem.h:25 */
  double L;
  float rdf_dr2;
  float* pos;      /* float[4N] */
  float* vel;
  float* frc;
  int rdf[256];
  double U, T, K, V, P, Pshear, t;
  double av_U_tot, av_T_tot, av_p_tot;
  int av_iters;
  int nthreads;    /* forces: 1 -> orc_forces_ref, >1 -> orc_forces_ref_omp */
} orc_sys;

/* allocate, L and dr2 rules (MDSystem.cpp:66-97); state zeroed */
int  orc_sys_init(orc_sys* s, int N, double T0, double rho, int canonical, int bc);
void orc_sys_free(orc_sys* s);
/* CalculateForces() CPU branch          MDSystem.cpp:232-311 */
void orc_sys_forces(orc_sys* s);
/* CalculateParameters() CPU branch      MDSystem.cpp:325-359 */
void orc_sys_params(orc_sys* s);
/* ApplyBoundaryConditions()             MDSystem.cpp:406-436 */
void orc_sys_apply_bc(orc_sys* s);
/* Integrate(dt): EVN or TVN, BCs, params, t += dt   MDSystem.cpp:438-510,578-582 */
void orc_sys_integrate(orc_sys* s, double dt);

/* Speed histogram of initvelo / updatevelo (MDSystem.cpp:651-694): counts
 * of bin (int)(sqrt((double)fl32(|v|^2)) / step) for bins < nbins. */
void orc_velocity_histogram(int N, const float* vel4, double step, int nbins, int64_t* counts);

/* One Integrate(dt) step (MDSystem.cpp:442-510, EVN or TVN, then the
 * boundary conditions) with every quantity in double, for the precision
 * study: x/v/f double[3N] (f holds the forces of x on entry and on exit),
 * forces from orc_forces_fp64d with eps, sigma. obs3 = {E_pot, E_kin, W}
 * after the step. */
void orc_fp64_step(int N, double* x, double* v, double* f, double L, int bc, int canonical,
                   double T0, double eps, double sigma, double dt, double* obs3, int nthreads);

/* ---------------- adjoint (FP64) ---------------- */

/* One adjoint pass in FP64 at double positions x3 with vector z3 (double[3N]):
 *   hz3 = H z (H = dF/dx),  f3 = F,  gw3 = dW/dx,
 *   scal8 = {E, W, dE/dsigma, dW/dsigma, (dF/dsigma)^T z, F^T z, 0, 0}.
 * Minimum image as orc_forces_fp64; self pair excluded by index. */
void orc_adjoint_pass_fp64(int N, const double* x3, const double* z3, double L, int periodic,
                           double eps, double sigma, double* hz3, double* f3, double* gw3,
                           double* scal8, int nthreads);

/* FP64 trajectory of M EVN velocity-Verlet steps from (x0, v0) (no wrap: the
 * minimum image makes forces wrap-invariant) and the loss
 *   L = (1/M) sum_{m=1..M} E(x_m) + w_p ((1/M) sum_m P_m - p_target)^2,
 *   P_m = (W(x_m)/3 + sum |v_m|^2 / 3) / L^3.
 * out3 = {L, mean E, mean P}. */
void orc_traj_loss_fp64(int N, const double* x0, const double* v0, double L, int periodic,
                        double eps, double sigma, double dt, int M, double p_target, double w_p,
                        double* out3, int nthreads);
/* Its gradient by the adjoint sweep over all stored states:
 * grad2 = {dL/deps, dL/dsigma}, dx0 = dL/dx0 (double[3N], may be NULL). */
void orc_traj_grad_fp64(int N, const double* x0, const double* v0, double L, int periodic,
                        double eps, double sigma, double dt, int M, double p_target, double w_p,
                        double* out3, double* grad2, double* dx0, int nthreads);

/* ---------------- fluctuation observables (lj_oracle_flucs.c) ---------------- */

/* alpha ladder of the batch counters: alpha = step, += step while alpha < limit
 * (run-fluctuations-aux.h:198 with limit 1 - 1e-9, :251 with 1 + 1e-9);
 * returns its size, writes up to cap values (alphas may be NULL) */
int orc_fl_ladder(double step, double limit, double* alphas, int cap);
/* GetNSubsystemBatch (run-fluctuations-aux.h:191-241): cumulative counts */
int orc_fl_nsub_batch(int N, const float* pos4, double L, double alpha_step, int type, int* cnts,
                      int cap);
/* GetNsubVzBatch (run-fluctuations-aux.h:246-278) */
int orc_fl_vz_batch(int N, const float* vel4, double vcut_max, double alpha_step, int type,
                    int* cnts, int cap);
/* GetNSubsystem (:128-160), GetNsubVz (:165-186) */
int orc_fl_nsub(int N, const float* pos4, double L, double fraction, int type);
int orc_fl_vz(int N, const float* vel4, double vcut, int type);
/* count of MDSystem::Fluctuations (MDSystem.cpp:708-719) */
long long orc_fl_closed_cube(int N, const float* pos4, double L, double fraction);

/* TimeAverage (time-average-aux.h:27-67) over NumberStatistics moment sums */
typedef struct orc_tavg {
  double sums[5]; /* sum of N^k, k = 0..4 */
  int64_t nobs;
  double last, corr;
  int iters;
} orc_tavg;
#define ORC_TAVG_NOUT 7
void orc_tavg_init(orc_tavg* a);
void orc_tavg_add(orc_tavg* a, double val);
/* {mean, variance, s, mean error * sqrt(s), scaled variance, scaled variance
 * error, plain mean error} */
void orc_tavg_eval(const orc_tavg* a, double* out);

#ifdef __cplusplus
}
#endif
#endif
