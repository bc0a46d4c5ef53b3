/*
 * lj_hip.h -- C-ABI of the MI355X (gfx950) Lennard-Jones time step.
 *
 * Three groups of entry points, all with C linkage and plain pointers/sizes:
 *
 *  (1) Compatibility set.  The exact functions that the reference's
 *      MDSystem.cpp declares under USE_CUDA_TOOLKIT (MDSystem.cpp:10-25) and
 *      MDSystem.cu defines (MDSystem.cu:155-299), same names, signatures and
 *      argument meaning.  Linking the reference's own MDSystem.cpp against
 *      libljhip.so instead of MDSystem.cu puts the HIP path behind its
 *      unchanged CPU/GPU switch (MDSystem.cpp:240).  They return void like
 *      the originals; a failure is reported on stderr and kept for
 *      lj_last_error() instead of exit()ing (set LJ_COMPAT_ABORT=1 to abort).
 *
 *  (2) Context API.  A context owns one box: device-resident positions,
 *      velocities and forces (AoS float4, the reference layout), scratch,
 *      one HIP stream and the captured step graph.  Every call returns
 *      LJ_OK (0) or a negative status and never exits the process.
 *      Calls are synchronous unless documented otherwise; distinct contexts
 *      may be used from distinct host threads (their work, copies and waits
 *      stay on their own streams, so one thread's step capture is not
 *      disturbed by another's calls; the compatibility set (1) keeps the
 *      reference's synchronous semantics instead).
 *
 *  (3) Device-pointer API for callers that keep their own device buffers
 *      (the PyTorch binding): forces and the adjoint pair kernel on caller
 *      pointers, enqueued on a caller stream.
 */
#ifndef LJ_HIP_H
#define LJ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status                                                              */
/* ------------------------------------------------------------------ */
#define LJ_OK 0
#define LJ_ERR_ARG (-1)         /* invalid argument (null, size, mode) */
#define LJ_ERR_DEVICE (-2)      /* no HIP device or a HIP runtime error */
#define LJ_ERR_ALLOC (-3)       /* device or host allocation failed */
#define LJ_ERR_STATE (-4)       /* call not valid in the current state */
#define LJ_ERR_UNSUPPORTED (-5) /* mode not available for this shape */

/* Message of the last failure on the calling thread ("" if none). */
const char* lj_last_error(void);
/* Number of visible HIP devices (0 when none; never fails). */
int lj_device_count(void);
/* Library version string. */
const char* lj_version(void);

/* ------------------------------------------------------------------ */
/* (1) compatibility set -- reference MDSystem.cpp:10-25              */
/* ------------------------------------------------------------------ */

/* two device arrays of 4*numBodies floats        MDSystem.cu:158-165 */
void allocateNBodyArrays(float* vel[2], int numBodies);
/* one device array of 4*number floats            MDSystem.cu:167-173 */
void allocateArray(float** dest, int number);
/*                                                MDSystem.cu:175-179 */
void deleteNBodyArrays(float* vel[2]);
/*                                                MDSystem.cu:181-184 */
void deleteArray(float* arr);
/* device array of `number` ints                  MDSystem.cu:186-192 */
void allocateIntArray(int** dest, int number);
/*                                                MDSystem.cu:194-197 */
void deleteIntArray(int* arr);
/* 16*numBodies bytes device->host; pbo ignored   MDSystem.cu:199-210 */
void copyArrayFromDevice(float* host, const float* device, unsigned int pbo, int numBodies);
/* 16*numBodies bytes host->device                MDSystem.cu:212-216 */
void copyArrayToDevice(float* device, const float* host, int numBodies);
/* OpenGL interop: no-ops, as in the reference    MDSystem.cu:218-226 */
void registerGLBufferObject(unsigned int pbo);
void unregisterGLBufferObject(unsigned int pbo);
/* device synchronise                             MDSystem.cu:228 */
void threadSync(void);
/*
 * Forces, energy, virial and RDF of numBodies particles      MDSystem.cu:230-291
 *   Pos    device float4[numBodies] (x,y,z,w; w ignored)
 *   Force  device float4[numBodies] out: (Fx,Fy,Fz, sum_j (r^-12 - r^-6))
 *          with F = 4 sum_j r (12 r^-14 - 6 r^-8)               MDSystem.cu:21-56,105-108
 *   host_pressure out: sum_i (4/6) sum_j r.f_ij (virial only)   MDSystem.cu:108,136
 *   Lperiodic  1: minimum image with box side host_L
 *   host_RDF   out: 256 counts of ordered pairs, bin floor(fl32(r^2/dr2)) clamped to 255
 *   pairs with r^2 <= 1e-6 are skipped, the self pair included MDSystem.cu:37
 *   blockSize  accepted for signature compatibility (tiling is chosen internally;
 *              the reference's per-blockSize RDF-merge stride bug, MDSystem.cu:149,
 *              is not reproduced); q unused as in the reference.
 * Synchronous: returns when Force, *host_pressure and host_RDF hold the results.
 * The virial sum is a fixed-order FP64 reduction (deterministic), rounded to float.
 */
void calculateNForces(float* Pos, float* Force, float* host_pressure, int numBodies,
                      float host_L, int Lperiodic, int* host_RDF, float host_dr2,
                      int blockSize, int q);
/*
 * calculateNForces without device allocation.  No counterpart in the
 * reference: it serves callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and synchronous behaviour as calculateNForces, with
 * the same kernels (bit-identical results), but every device scratch buffer
 * lives in a file-scope device array of 12.3 MiB that each GPU the library
 * runs on holds until the process ends; no hipMalloc, hipFree or stream
 * creation.  Runs on the null stream.  numBodies <= 65536.  Returns LJ_OK or
 * a negative status (message in lj_last_error()).  The array is shared by
 * every call on a GPU and keeps the last call's contents: calls overlapping
 * on one GPU would race, so the calls of one process are serialised.
 */
int calculateNForcesNoAlloc(float* Pos, float* Force, float* host_pressure, int numBodies,
                            float host_L, int Lperiodic, int* host_RDF, float host_dr2,
                            int blockSize, int q);
/* Releases the compatibility path's pooled scratch.  Unlike cudaThreadExit
 * (MDSystem.cu:294-297) it does not reset the device, so other live arrays and
 * contexts stay valid. */
void threadExit(void);

/* ------------------------------------------------------------------ */
/* (2) context API                                                     */
/* ------------------------------------------------------------------ */

typedef struct lj_ctx lj_ctx;

enum { LJ_BC_PERIODIC = 0, LJ_BC_WALL = 1, LJ_BC_NONE = 2 }; /* MDSystem.h:24-25 */
enum { LJ_EVN = 0, LJ_TVN = 1 };                              /* MDSystem.h:20-21 */
enum {
  LJ_RDF_REF_CPU = 0, /* reference CPU rule: r^2 as MDSystem.cpp:269-279 (double min image),
                         bin = floor(r^2/dr2) exact, bins >= nbins dropped, self by index */
  LJ_RDF_REF_GPU = 1, /* reference GPU rule, bit pattern assumed (nvcc contraction),
                         bounded by the CPU rule: MDSystem.cu:26-42 (float lround min
                         image, r^2 = fma(z,z,fma(y,y,x*x)) -- the contraction nvcc
                         applies by default, assumed, since the reference's GPU build
                         cannot run here -- floor(fl32(r^2/dr2)), clamp to nbins-1,
                         pairs with r^2 <= 1e-6 skipped).  Its counts differ from
                         LJ_RDF_REF_CPU's only at bin edges and in the clamp bin. */
  LJ_RDF_RBINS = 2    /* r^2 as LJ_RDF_REF_CPU, nbins uniform in r on [0, rmax):
                         bin = max k with edges2[k] <= r^2, edges2[k] =
                         (float)((k*(rmax/nbins))^2) in double; r^2 >= edges2[nbins] dropped */
};
enum {
  LJ_PREC_FP32 = 0,        /* FP32 pair math and per-slice sums, FP64 cross-slice sums */
  LJ_PREC_COMPENSATED = 1, /* + per-particle sums in FP32 blocks of 32 pairs added by
                              TwoSum, inside each partial entry (slice or group pair) */
  LJ_PREC_FP64 = 2         /* double state, double pair math, the integrator formulas of
                              MDSystem.cpp:442-510 in double: the reference of the
                              precision study.  The float arrays are rounded mirrors
                              of the double state; lj_set/get_array_f64 reach it. */
};
enum { LJ_ARRAY_POS = 0, LJ_ARRAY_VEL = 1, LJ_ARRAY_FORCE = 2 }; /* MDSystem.cpp:586-630 */

typedef struct lj_config {
  int N;          /* particles                               MDSystem.h:12 */
  double T0;      /* initial / TVN temperature               MDSystem.h:15 */
  double rho;     /* number density                          MDSystem.h:18 */
  int canonical;  /* LJ_EVN or LJ_TVN                        MDSystem.h:21 */
  int boundary;   /* LJ_BC_*                                 MDSystem.h:25 */
  double eps;     /* LJ well depth (reference: 1)            */
  double sigma;   /* LJ diameter (reference: 1)              */
  int precision;  /* LJ_PREC_*                               */
  int jslices;    /* j-range split of the force kernel; 0 = automatic */
  int device;     /* HIP device ordinal                      */
} lj_config;

/* Observables of the last evaluated state.  Field meaning as the reference's
 * public members MDSystem.h:61-66,79-80,97-99 (CalculateParameters,
 * MDSystem.cpp:325-359): P includes the kinetic term and the volume, V is the
 * potential energy.  W is the virial sum_{i<j} r_ij.f_ij (P = (W/3 + N T)/vol).
 * Pshear is not computed on this path (nor on the reference's GPU branch). */
typedef struct lj_obs {
  double U, T, K, V, P, Pshear, t, W;
  double av_U_tot, av_T_tot, av_p_tot;
  int64_t av_iters;
  int64_t step;
} lj_obs;

/* defaults of MDSystemConfiguration (MDSystem.h:32-40) plus eps = sigma = 1 */
void lj_config_init(lj_config* cfg);
int lj_create(const lj_config* cfg, lj_ctx** out);
int lj_destroy(lj_ctx* ctx);
/* box side L = (N/rho)^(1/3) (MDSystem.cpp:70) */
double lj_box_length(const lj_ctx* ctx);
/* reference RDF bin width in r^2 (MDSystem.cpp:93-95) */
float lj_rdf_dr2_default(int N);

/* Initial state of the benchmark configurations: simple-cubic lattice
 * (MDSystem.cpp:150-168), Gaussian velocities from SplitMix64(seed) with
 * Box-Muller, momentum removed and rescaled to exactly T0
 * (MDSystem.cpp:179-180).  Computed on the host, uploaded; then forces and
 * observables are evaluated as MDSystem::Reinitialize does (MDSystem.cpp:101-103). */
int lj_sample_initial_state(lj_ctx* ctx, uint64_t seed);
/* The same with a choice of velocity distribution:
 *   LJ_INIT_GAUSSIAN  as lj_sample_initial_state;
 *   LJ_INIT_MAXWELL   the reference's own sampler (MDSystem.cpp:36-53,170-177):
 *                     Maxwell speeds sqrt(2 T0) t, t drawn by the reference's
 *                     rejection rule (density ~ t^2 exp(-t^2)), directions
 *                     with cos(theta) and phi uniform, from two SplitMix64
 *                     streams of the seed instead of the clock-seeded MTRand;
 * then the momentum is removed and T rescaled to exactly T0 (MDSystem.cpp:179-180). */
enum { LJ_INIT_GAUSSIAN = 0, LJ_INIT_MAXWELL = 1 };
int lj_sample_initial_state_ex(lj_ctx* ctx, uint64_t seed, int velocities);

/* host float[4N] <-> device state           MDSystem.cpp:586-630 */
int lj_set_array(lj_ctx* ctx, int type, const float* host4N);
int lj_get_array(lj_ctx* ctx, int type, float* host4N);
/* host double[4N] <-> device state.  In LJ_PREC_FP64 these reach the double
 * state itself (set also refreshes the float mirror); in the FP32 modes the
 * values are rounded to / widened from the float arrays. */
int lj_set_array_f64(lj_ctx* ctx, int type, const double* host4N);
int lj_get_array_f64(lj_ctx* ctx, int type, double* host4N);

/* CalculateForces (MDSystem.cpp:232-311): device forces of the current
 * positions, E_pot and virial into the observables (out may be NULL). */
int lj_compute_forces(lj_ctx* ctx, lj_obs* out);
/* CalculateParameters (MDSystem.cpp:325-359) from the device state. */
int lj_compute_parameters(lj_ctx* ctx, lj_obs* out);
int lj_get_observables(lj_ctx* ctx, lj_obs* out);

/* nsteps of Integrate(dt) (MDSystem.cpp:438-583) fully on the device:
 * velocity Verlet EVN or TVN, boundary conditions, observables; captured
 * once into a HIP graph and replayed.  hist (may be NULL) receives 5 doubles
 * per step {U, T, K, V, P}.  When the context runs the pair-once force kernel
 * (lj_force_plan_info), the step's force pass sums the potential energy and
 * virial per workgroup instead of per particle: V, W and P are unchanged, but
 * the force array's w (per-particle energy, read by no caller of the
 * reference) is left 0 until the next lj_compute_forces. */
int lj_integrate(lj_ctx* ctx, double dt, int nsteps, double* hist);

/* Enqueue nsteps of the captured step graph and return without waiting (no
 * history copy).  lj_synchronize() waits; observables are read afterwards. */
int lj_integrate_async(lj_ctx* ctx, double dt, int nsteps);

/* Histogram of ordered-pair distances of the current positions.
 * counts: host uint64[nbins]; accumulate != 0 adds to the device histogram
 * kept from earlier calls with the same rule/nbins, 0 clears it first.
 * dr2 <= 0 picks the reference rule; rmax <= 0 picks L/2. */
int lj_rdf(lj_ctx* ctx, int rule, int nbins, float dr2, double rmax, uint64_t* counts,
           int accumulate);

/* Average time of one RDF pass over the current positions (HIP events,
 * microseconds), arguments as lj_rdf; the device histogram is left holding
 * the timed passes, so the next lj_rdf call starts from zero. */
int lj_time_rdf(lj_ctx* ctx, int rule, int nbins, float dr2, double rmax, int reps, double* us);

/* Speed histogram of the current velocities as MDSystem::initvelo /
 * updatevelo bin them (MDSystem.cpp:651-694): counts[b] = number of particles
 * with (int)(sqrt((double)fl32(vx*vx + vy*vy + vz*vz)) / step) == b, for
 * b < nbins (the reference's maxind = (int)(vmax / step) + 1); nbins <= 16384. */
int lj_velocity_histogram(lj_ctx* ctx, double step, int nbins, uint64_t* counts);

/* One GPU's share of a force pass sharded by i-particles over nshards GPUs
 * (no counterpart in the reference, which runs on one device; the single-GPU
 * part of the multi-GPU design).  The share is the forces on particles
 * [i0, i1), whole 64-particle tiles: with T = ceil(N/64),
 * i0 = 64*floor(shard*T/nshards), i1 = min(N, 64*floor((shard+1)*T/nshards)),
 * 1 <= nshards <= T, from all N current positions, on the sliced kernel with
 * jslices j slices (0: enough slices to fill the device at that share).  The
 * rows go to rows i0..i1-1 of d_force4 (float4[N] on the device; NULL: the
 * context's force array, which the next lj_integrate's first drift reads, so
 * run every share before stepping), and the share's E_pot and W partial sums
 * to ev[0], ev[1] (may be NULL): the all-reduce's inputs.  With the jslices
 * of a full sliced evaluation (lj_config.jslices) the rows are bit-identical
 * to it.  FP32 and compensated modes. */
int lj_forces_shard(lj_ctx* ctx, int shard, int nshards, int jslices, float* d_force4,
                    double* ev);
/* the share's force kernel alone, averaged over reps launches (HIP events) */
int lj_time_forces_shard(lj_ctx* ctx, int shard, int nshards, int jslices, int reps, double* us);
/* The other way to shard (no counterpart in the reference): the time step's
 * pair-once pass with its group-pair items dealt out, share `shard` of
 * nshards taking the contiguous items [shard*M/nshards, (shard+1)*M/nshards)
 * of the M workgroup items (every item carries the same work).  Each item
 * writes its own partial entries, so across GPUs an all-reduce of the partial
 * buffers would give the full pass's.  Times the share's launch (HIP events,
 * reps launches; the context's partial buffers are overwritten).  FP32 mode,
 * pair-once pass with paired diagonal items (N >= about 11000). */
int lj_time_force_items_share(lj_ctx* ctx, int shard, int nshards, int reps, double* us);

/* Shear stress of the current state as the reference's CPU path computes it
 * (MDSystem.cpp:299,309,335,353): (sum over ordered pairs of -r_x f_y / 2 +
 * sum_i fl32(-v_x v_y)) / (N / rho), the pair terms in the force pass's FP32
 * arithmetic with FP64 sums.  The reference's GPU branch does not compute it,
 * and neither does the time step (lj_obs.Pshear stays 0); this runs one
 * all-pairs pass on demand.  FP64 contexts use the float mirrors of the state. */
int lj_shear_pressure(lj_ctx* ctx, double* pshear);

int lj_renormalize_velocities(lj_ctx* ctx, int recompute_tkin); /* MDSystem.cpp:375-389 */
int lj_renormalize_to_energy(lj_ctx* ctx, double ust);          /* MDSystem.cpp:391-404 */
int lj_reset_averaging(lj_ctx* ctx);                            /* MDSystem.cpp:701-705 */
int lj_apply_boundary_conditions(lj_ctx* ctx);                  /* MDSystem.cpp:406-436 */
int lj_set_mode(lj_ctx* ctx, int canonical, int boundary);      /* MDSystem.h:161-163 */
int lj_synchronize(lj_ctx* ctx);

/* device pointers of the state arrays (float4[N]) for zero-copy callers */
int lj_device_arrays(lj_ctx* ctx, float** pos, float** vel, float** force);

/* time of the force kernel alone, averaged over `reps` launches on the
 * context stream, measured with HIP events (for roofline reporting). */
int lj_time_force_kernel(lj_ctx* ctx, int reps, double* us_per_launch);
/* the same launches (1 <= reps <= 4096, FP32 mode), each stamped with its
 * own start and stop events (hipExtLaunchKernel): us_kernel = the mean of the
 * launches' own durations (what a kernel trace reports), us_span = events
 * around all of them / reps (the back-to-back rate, dispatch gaps included) */
int lj_time_force_kernel_stamped(lj_ctx* ctx, int reps, double* us_kernel, double* us_span);
/* which force kernel the context runs: pair_once = 1 for the kernel that
 * evaluates each unordered pair once (FP32 mode, N >= 3072, j slices
 * automatic), 0 for the sliced kernel; nslots = partial sums per particle */
int lj_force_plan_info(lj_ctx* ctx, int* pair_once, int* nslots);

/* ---- fluctuation observables: particle counts of sub-volumes and velocity
 * windows, counted and accumulated on the device ---- */
enum {
  LJ_FL_SLAB_X = 0, /* x / L below alpha      run-fluctuations-aux.h:223-234 (types 0-2) */
  LJ_FL_SLAB_Y = 1,
  LJ_FL_SLAB_Z = 2,
  LJ_FL_CUBE = 3,   /* centred open cube of volume fraction alpha              :209-222 */
  LJ_FL_VEL_X = 4,  /* |v_x| / vmax below alpha                                 :246-278 */
  LJ_FL_VEL_Y = 5,
  LJ_FL_VEL_Z = 6,
  LJ_FL_CUBE_CLOSED = 7 /* closed centred cube, MDSystem::Fluctuations MDSystem.cpp:708-719 */
};
/* One counter.  single = 0: the reference's ladder of windows alpha = step,
 * 2 step, ... accumulated in double below 1 - 1e-9 (slabs, cubes;
 * GetNSubsystemBatch, run-fluctuations-aux.h:191-241) or 1 + 1e-9 (velocity
 * windows with vcut_max = vmax; GetNsubVzBatch, :246-278), one entry per
 * window with the cumulative count.  single = 1: one window, alpha the volume
 * fraction (GetNSubsystem :128-160; MDSystem::Fluctuations) or, for velocity
 * windows, |v| < vmax (GetNsubVz :165-186).  Counts use the reference's double
 * arithmetic on the float state and match it exactly. */
typedef struct lj_fl_ladder {
  int type;     /* LJ_FL_* */
  int single;   /* 0: ladder of step alpha; 1: one window */
  double alpha; /* ladder step, or the volume fraction of one window */
  double vmax;  /* velocity types: vcut_max (ladder) or vcut (one window) */
} lj_fl_ladder;
/* statistics of an entry, LJ_FL_NSTAT doubles: sum n^0 .. sum n^4 (as the
 * moment sums of NumberStatistics), sum over samples t >= 2 of n_t n_(t-1),
 * the last n, the number of samples (TimeAverage, time-average-aux.h:27-48) */
#define LJ_FL_NSTAT 8
/* Sets the counters of the context (n = 0 removes them) and clears their
 * statistics.  At most 64 ladders and 4096 entries. */
int lj_fl_configure(lj_ctx* ctx, const lj_fl_ladder* ladders, int n);
/* entries of the configured counters (the ladders' entries in order) */
int lj_fl_num_entries(lj_ctx* ctx, int* entries);
/* per entry: its alpha (ladders: the accumulated step; one window: the
 * fraction, or vcut for a velocity window) */
int lj_fl_values(lj_ctx* ctx, double* values);
/* counts the current state and adds the counts to the statistics */
int lj_fl_sample(lj_ctx* ctx);
/* the last sample's cumulative counts, per entry */
int lj_fl_counts(lj_ctx* ctx, int64_t* counts);
/* accumulated statistics, LJ_FL_NSTAT per entry */
int lj_fl_stats(lj_ctx* ctx, double* stats);
int lj_fl_reset(lj_ctx* ctx);

/* Sampling inside the time step (captured into the step graphs).
 *   rdf = LJ_SAMPLE_RDF_LAST: every step bins the positions its force pass
 *         sees, replacing the previous step's histogram, as the reference
 *         refills NdNdr2 in every CalculateForces (MDSystem.cpp:254,281-285);
 *   rdf = LJ_SAMPLE_RDF_SUM: the steps' histograms are added up;
 *   rdf_rule, rdf_nbins, rdf_dr2, rdf_rmax as lj_rdf (FP32 modes only);
 *   fluctuations = 1: lj_fl_sample after every step, on the state the step
 *         leaves (what a caller reads after Integrate).
 * Steps that sample run without fusing the next step's drift into the end of
 * the step, so the sampled state is the step's final one. */
enum { LJ_SAMPLE_RDF_OFF = 0, LJ_SAMPLE_RDF_LAST = 1, LJ_SAMPLE_RDF_SUM = 2 };
typedef struct lj_step_sampling {
  int rdf;
  int rdf_rule;
  int rdf_nbins;
  float rdf_dr2;
  double rdf_rmax;
  int fluctuations;
} lj_step_sampling;
int lj_set_step_sampling(lj_ctx* ctx, const lj_step_sampling* s);
/* the step-sampled RDF histogram (uint64[rdf_nbins]) and the number of steps
 * it holds (1 for LJ_SAMPLE_RDF_LAST once a step has run); clear != 0 zeroes
 * it afterwards */
int lj_step_rdf_counts(lj_ctx* ctx, uint64_t* counts, int64_t* nsamples, int clear);

/* ------------------------------------------------------------------ */
/* (3) device-pointer API                                              */
/* ------------------------------------------------------------------ */

/* Forces of caller-owned positions d_pos4 (float4[N], device) into d_force4
 * (float4[N], w = per-particle sum ((s/r)^12-(s/r)^6)); d_ev (device, may be
 * NULL) receives {E_pot, W}.  Box, boundary, eps, sigma from ctx; N must equal
 * the context's N.  Enqueued on `stream` (NULL = context stream), async.
 * Always the FP32 pair pass (the compensated one in LJ_PREC_COMPENSATED),
 * also in an LJ_PREC_FP64 context. */
int lj_forces_device(lj_ctx* ctx, const float* d_pos4, float* d_force4, double* d_ev,
                     void* stream);

/* Device-to-device state transfer (float4[N]) on `stream` (NULL = context stream). */
int lj_set_array_device(lj_ctx* ctx, int type, const float* d_src4, void* stream);
int lj_get_array_device(lj_ctx* ctx, int type, float* d_dst4, void* stream);

/* LJ parameters of the context (the reference uses eps = sigma = 1). */
int lj_set_lj_params(lj_ctx* ctx, double eps, double sigma);

/* ---- adjoint (differentiable) path ---- */
/* One adjoint pair pass at positions d_pos4 with vector d_z4 (float4[N], w ignored):
 *   d_hz4  (double4[N])  (H z)_i, H = dF/dx (Hessian-vector product of the pair force)
 *   d_f4   (double4[N])  forces, w = sum_j ((s/r)^12 - (s/r)^6)
 *   d_gw4  (double4[N])  dW/dx_i of the virial W = sum_{i<j} r.f
 *   d_scal8 (double[8])  {E, W, dE/dsigma, dW/dsigma, (dF/dsigma)^T z, F^T z, 0, 0}
 * Any output may be NULL.  Synchronous. */
int lj_adjoint_pass_device(lj_ctx* ctx, const float* d_pos4, const float* d_z4, double* d_hz4,
                           double* d_f4, double* d_gw4, double* d_scal8, void* stream);

/* Loss of nsteps EVN velocity-Verlet steps (MDSystem.cpp:442-464) from the
 * context state:  L = (1/M) sum_m E_pot(x_m) + w_p ((1/M) sum_m P_m - p_target)^2
 * over the states after steps m = 1..M, P as CalculateParameters (MDSystem.cpp:348-350).
 * out3 = {L, mean E_pot, mean P}.  The context state is left unchanged. */
int lj_trajectory_loss(lj_ctx* ctx, double dt, int nsteps, double p_target, double w_p,
                       double* out3);
/* The same loss and its gradient with respect to eps, sigma (grad_eps_sigma[2])
 * and the initial positions (dx0_host4: host float4[N], and/or d_dx0_4: device
 * float4[N]; either may be NULL), by a reverse sweep with one Hessian-vector
 * product per state, checkpoints every `checkpoint_every` steps and
 * recomputation inside each segment (checkpoint_every <= 0: one segment).
 * times_ms3 (may be NULL) = {forward, recompute, reverse} milliseconds.
 * EVN only; periodic or open boxes.  The context state is left unchanged. */
int lj_trajectory_grad(lj_ctx* ctx, double dt, int nsteps, int checkpoint_every, double p_target,
                       double w_p, double* out3, double* grad_eps_sigma, float* dx0_host4,
                       float* d_dx0_4, double* times_ms3);

/* ------------------------------------------------------------------ */
/* (4) batched ensemble: many independent boxes in one launch          */
/* ------------------------------------------------------------------ */
/* Every box runs the single-box time step above (reference
 * MDSystem::Integrate per box) with its own N, density and temperature, in
 * one set of launches per step: the force pass is one persistent launch that
 * drains a work queue of (box, i-tile, j-slice) items; the per-particle
 * kernels use workgroups that never straddle boxes; the whole step is
 * captured once into a HIP graph.  Per-box energy and sub-volume particle
 * counts are accumulated on the device (the fluctuation observables of
 * src/tasks/run-fluctuations/include/run-fluctuations-aux.h:128-241 and
 * MDSystem::Fluctuations, MDSystem.cpp:708-730). */

typedef struct lj_batch lj_batch;

/* Tiling of the batched force pass (picked by lj_batch_autotune):
 *   ipl            i particles per lane (1 or 2): i-tile = 64*ipl particles
 *   jchunk         j particles per work item (waves per box = tiles x slices)
 *   blocks_per_cu  persistent grid of blocks_per_cu workgroups per CU pulling
 *                  items from a queue; 0 = one wave per item (no queue)
 *   pair_once      1: each unordered pair once, items = pairs of 256-particle
 *                  groups of a box, run by 4-wave workgroups: one per item
 *                  (blocks_per_cu 0) or a persistent grid of blocks_per_cu
 *                  workgroups per CU draining the item queue (ipl and jchunk
 *                  are then unused); 0: the tiling above */
typedef struct lj_batch_tuning {
  int ipl;
  int jchunk;
  int blocks_per_cu;
  int pair_once;
} lj_batch_tuning;

/* per-box statistics, LJ_BATCH_NSTAT doubles per box, accumulated over steps:
 * samples, sum U, sum U^2, sum V, sum V^2, sum K, sum K^2, sum P, sum P^2,
 * sum n, sum n^2, sum n^3, sum n^4 (n = particles in the centred sub-cube
 * of volume fraction f, MDSystem.cpp:711-723) */
#define LJ_BATCH_NSTAT 13

int lj_batch_create(int nboxes, const int* N, const double* rho, const double* T0, int canonical,
                    int boundary, int device, lj_batch** out);
int lj_batch_destroy(lj_batch* b);
int lj_batch_num_particles(const lj_batch* b);
/* lattice + seeded Gaussian state per box (box k uses seed + k), then forces */
int lj_batch_sample_initial_state(lj_batch* b, uint64_t seed);
/* box = -1: all boxes (flat float4[sum N]); else float4[N_box] */
int lj_batch_set_array(lj_batch* b, int box, int type, const float* host4N);
int lj_batch_get_array(lj_batch* b, int box, int type, float* host4N);
int lj_batch_compute_forces(lj_batch* b);
/* nsteps of the captured batched step */
int lj_batch_integrate(lj_batch* b, double dt, int nsteps);
/* the same with the per-step history: hist (host, nsteps * nboxes * 5
 * doubles, step-major) receives {U, T, K, V, P} of every box after every step */
int lj_batch_integrate_hist(lj_batch* b, double dt, int nsteps, double* hist);
int lj_batch_get_observables(lj_batch* b, lj_obs* out /* nboxes */);
int lj_batch_set_subvolume(lj_batch* b, double fraction);
int lj_batch_get_stats(lj_batch* b, double* out /* nboxes * LJ_BATCH_NSTAT */);
int lj_batch_reset_stats(lj_batch* b);
int lj_batch_set_tuning(lj_batch* b, const lj_batch_tuning* t);
int lj_batch_get_tuning(lj_batch* b, lj_batch_tuning* t);
/* average time of one batched force pass (HIP events), microseconds: the
 * time step's pass (with pair_once: force-only partials and per-box energy and
 * virial totals; after lj_batch_integrate with pair_once the force array's w
 * is 0 until the next lj_batch_compute_forces) */
int lj_batch_time_force(lj_batch* b, int reps, double* us);
/* Times every candidate tiling on this batch (reps force passes each), keeps
 * the fastest; cached in the file named by LJ_AUTOTUNE_CACHE when set. */
int lj_batch_autotune(lj_batch* b, int reps, lj_batch_tuning* best, double* best_us);

/* Fluctuation counters of every box (the single-box lj_fl_* for each box with
 * its own side L): E entries per box, outputs box-major (nboxes x E). */
int lj_batch_fl_configure(lj_batch* b, const lj_fl_ladder* ladders, int n);
int lj_batch_fl_num_entries(lj_batch* b, int* entries);
int lj_batch_fl_sample(lj_batch* b);
int lj_batch_fl_counts(lj_batch* b, int64_t* counts /* nboxes * E */);
int lj_batch_fl_stats(lj_batch* b, double* stats /* nboxes * E * LJ_FL_NSTAT */);
int lj_batch_fl_reset(lj_batch* b);
/* fluctuations = 1: the counters sample every box after every step of
 * lj_batch_integrate (inside the captured step) */
int lj_batch_set_step_sampling(lj_batch* b, int fluctuations);

#ifdef __cplusplus
}
#endif
#endif /* LJ_HIP_H */
