/*
 * lj_oracle.c -- CPU restatement of the reference Lennard-Jones time step.
 *
 * TEST INFRASTRUCTURE ONLY (the test oracle).  Not linked into the product.
 * Reference: vlvovch/lennard-jones-cuda, src/library/MDSystem.cpp (CPU path)
 * and src/library/MDSystem.cu (GPU-variant semantics for the RDF rule).
 * Build: -O2 -ffp-contract=off, no -ffast-math (the reference's x86-64 build
 * has no FMA contraction, so neither may this one).
 */
#include "lj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793238462643

/* ---------------- initial state ---------------- */

int orc_lattice_n(int N) { return (int)ceil(pow((double)N, 1. / 3.)); } /* MDSystem.cpp:150 */

double orc_box_length(int N, double rho) { return pow(N / rho, 1. / 3.); } /* MDSystem.cpp:70 */

float orc_rdf_dr2(int N) /* MDSystem.cpp:93-95 */
{
  double a = 0.2 * sqrt(100. / N);
  float dr2 = (float)(a > 0.05 ? a : 0.05);
  if (250 * dr2 < 25.0) dr2 = (float)(25.0 / 250);
  return dr2;
}

void orc_lattice_positions(int N, double L, float* pos4) /* MDSystem.cpp:150-168 */
{
  int n = orc_lattice_n(N);
  double dL = L / n;
  for (int k = 0; k < N; ++k) {
    int ix = k % n, iy = (k / n) % n, iz = k / (n * n);
    pos4[4 * k + 0] = (float)((ix + 0.5) * dL);
    pos4[4 * k + 1] = (float)((iy + 0.5) * dL);
    pos4[4 * k + 2] = (float)((iz + 0.5) * dL);
    pos4[4 * k + 3] = (float)(L / 150.f);
  }
}

double orc_splitmix_uniform(uint64_t* state)
{
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void orc_gaussian_velocities(int N, double T0, uint64_t seed, float* vel4)
{
  /* Box-Muller pairs over the flat component index m = 3k + c. */
  uint64_t st = seed;
  double sT = sqrt(T0);
  int M = 3 * N;
  for (int m = 0; m < M; m += 2) {
    double u1 = 1.0 - orc_splitmix_uniform(&st); /* (0, 1] */
    double u2 = orc_splitmix_uniform(&st);
    double rad = sqrt(-2.0 * log(u1));
    double g0 = rad * cos(2.0 * ORC_PI * u2);
    double g1 = rad * sin(2.0 * ORC_PI * u2);
    vel4[4 * (m / 3) + m % 3] = (float)(sT * g0);
    if (m + 1 < M) vel4[4 * ((m + 1) / 3) + (m + 1) % 3] = (float)(sT * g1);
  }
  for (int k = 0; k < N; ++k) vel4[4 * k + 3] = 0.f;
  orc_correct_momentum(N, vel4);                                     /* :179 */
  orc_renormalize_velocities(N, T0, orc_kinetic_temperature(N, vel4), vel4); /* :180 */
}

void orc_correct_momentum(int N, float* v) /* MDSystem.cpp:183-216 */
{
  double tPx = 0., tPy = 0., tPz = 0., totmass = 0.;
  for (int i = 0; i < 4 * N; i += 4) {
    tPx += v[i];
    tPy += v[i + 1];
    tPz += v[i + 2];
    totmass += 1.;
  }
  for (int i = 0; i < 4 * N; i += 4) {
    v[i] = (float)(v[i] + -tPx / totmass);
    v[i + 1] = (float)(v[i + 1] + -tPy / totmass);
    v[i + 2] = (float)(v[i + 2] + -tPz / totmass);
  }
}

double orc_kinetic_temperature(int N, const float* v) /* MDSystem.cpp:361-373 */
{
  double ret = 0.;
  for (int i = 0; i < 4 * N; i += 4) {
    float s = v[i] * v[i] + v[i + 1] * v[i + 1] + v[i + 2] * v[i + 2];
    ret += s;
  }
  ret *= 1. / 3. / N;
  return ret;
}

void orc_renormalize_velocities(int N, double T0, double Tkin, float* v) /* :375-389 */
{
  for (int i = 0; i < 4 * N; i += 4) {
    v[i] = (float)(v[i] * sqrt(T0 / Tkin));
    v[i + 1] = (float)(v[i + 1] * sqrt(T0 / Tkin));
    v[i + 2] = (float)(v[i + 2] * sqrt(T0 / Tkin));
  }
}

int orc_fast_round(float x) /* MDSystem.cpp:732-739 */
{
  if (x > 0) return (int)(x + 0.5f);
  return (int)(x - 0.5f);
}

/* ---------------- forces ---------------- */

/* the min-image step of MDSystem.cpp:273-277: float r, double L */
static inline float ref_minimage(float r, double L)
{
  return (float)(r - L * orc_fast_round((float)(r / L)));
}

/* one reference pair for particle i (MDSystem.cpp:269-299), accumulates into
 * float force f[3] and double V/P/Pshear */
static inline void ref_pair(const float* pi, const float* pj, double L, int bc,
                            float dr2, float* f, double* V, double* P,
                            double* Psh, int* rdf256)
{
  float rx = pi[0] - pj[0];
  float ry = pi[1] - pj[1];
  float rz = pi[2] - pj[2];
  if (bc == 0) {
    rx = ref_minimage(rx, L);
    ry = ref_minimage(ry, L);
    rz = ref_minimage(rz, L);
  }
  float r2f = rx * rx + ry * ry + rz * rz; /* float expression (:279) */
  double r2 = r2f;
  {
    int ind = (int)floor(r2 / dr2); /* :282-284 */
    if ((unsigned)ind < 256u) rdf256[ind]++;
  }
  r2 = 1 / r2;                       /* :289 */
  double r6 = r2 * r2 * r2;
  double fijx = r2 * (12 * r6 * r6 - 6 * r6) * rx;
  double fijy = r2 * (12 * r6 * r6 - 6 * r6) * ry;
  double fijz = r2 * (12 * r6 * r6 - 6 * r6) * rz;
  f[0] = (float)(f[0] + fijx);       /* :294-296, float accumulation */
  f[1] = (float)(f[1] + fijy);
  f[2] = (float)(f[2] + fijz);
  *V += (1 * r6 * r6 - 1 * r6);      /* :297 */
  *P += rx * fijx + ry * fijy + rz * fijz; /* :298 */
  *Psh += -rx * fijy;                /* :299 */
}

void orc_forces_ref(int N, const float* pos, double L, int bc, float dr2,
                    float* frc, double* V_, double* P_, double* Psh_, int* rdf)
{
  memset(rdf, 0, 256 * sizeof(int)); /* :254 */
  double V = 0., P = 0., Psh = 0.;
  for (int i = 0; i < 4 * N; i += 4) {
    float f[3] = {0.f, 0.f, 0.f};
    for (int j = 0; j < 4 * N; j += 4)
      if (j != i) ref_pair(pos + i, pos + j, L, bc, dr2, f, &V, &P, &Psh, rdf);
    frc[i] = f[0] * 4; /* :303-305 */
    frc[i + 1] = f[1] * 4;
    frc[i + 2] = f[2] * 4;
  }
  P *= 4. / 3. / 2.; /* :307-309 */
  V *= 4. / 2.;
  Psh *= 4. / 2.;
  *V_ = V;
  *P_ = P;
  *Psh_ = Psh;
}

void orc_forces_ref_omp(int N, const float* pos, double L, int bc, float dr2,
                        float* frc, double* V_, double* P_, double* Psh_,
                        int* rdf, int nthreads)
{
  double* part = (double*)calloc((size_t)3 * N, sizeof(double));
  memset(rdf, 0, 256 * sizeof(int));
#pragma omp parallel num_threads(nthreads)
  {
    int lrdf[256];
    memset(lrdf, 0, sizeof(lrdf));
#pragma omp for schedule(dynamic, 16)
    for (int ii = 0; ii < N; ++ii) {
      int i = 4 * ii;
      float f[3] = {0.f, 0.f, 0.f};
      double V = 0., P = 0., Psh = 0.;
      for (int j = 0; j < 4 * N; j += 4)
        if (j != i) ref_pair(pos + i, pos + j, L, bc, dr2, f, &V, &P, &Psh, lrdf);
      frc[i] = f[0] * 4;
      frc[i + 1] = f[1] * 4;
      frc[i + 2] = f[2] * 4;
      part[3 * ii] = V;
      part[3 * ii + 1] = P;
      part[3 * ii + 2] = Psh;
    }
#pragma omp critical
    for (int b = 0; b < 256; ++b) rdf[b] += lrdf[b];
  }
  double V = 0., P = 0., Psh = 0.;
  for (int ii = 0; ii < N; ++ii) {
    V += part[3 * ii];
    P += part[3 * ii + 1];
    Psh += part[3 * ii + 2];
  }
  free(part);
  *V_ = V * (4. / 2.);
  *P_ = P * (4. / 3. / 2.);
  *Psh_ = Psh * (4. / 2.);
}

static void fp64_forces_core(int N, const double* x, double L, int periodic,
                             double eps, double sigma, double skip_r2,
                             double* f3, double* ew, double* E, double* W,
                             int nthreads)
{
  double s2 = sigma * sigma;
  double Esum = 0., Wsum = 0.;
  if (nthreads <= 0) {
#ifdef _OPENMP
    nthreads = omp_get_num_procs();
#else
    nthreads = 1;
#endif
  }
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads) reduction(+ : Esum, Wsum)
  for (int i = 0; i < N; ++i) {
    double fx = 0., fy = 0., fz = 0., e = 0., w = 0.;
    for (int j = 0; j < N; ++j) {
      if (skip_r2 <= 0. && j == i) continue;
      double dx = x[3 * i] - x[3 * j];
      double dy = x[3 * i + 1] - x[3 * j + 1];
      double dz = x[3 * i + 2] - x[3 * j + 2];
      if (periodic) { /* half away from zero, as fast_round / lround (MDSystem.cpp:732, .cu:31) */
        dx -= L * round(dx / L);
        dy -= L * round(dy / L);
        dz -= L * round(dz / L);
      }
      double r2 = dx * dx + dy * dy + dz * dz;
      if (skip_r2 > 0. && !(r2 > skip_r2)) continue;
      double q2 = s2 / r2;
      double q6 = q2 * q2 * q2;
      double q12 = q6 * q6;
      double s = 24. * eps * (2. * q12 - q6) / r2; /* |f| / r */
      fx += s * dx;
      fy += s * dy;
      fz += s * dz;
      e += q12 - q6;
      w += s * r2;
    }
    if (f3) {
      f3[3 * i] = fx;
      f3[3 * i + 1] = fy;
      f3[3 * i + 2] = fz;
    }
    if (ew) {
      ew[2 * i] = e;
      ew[2 * i + 1] = w;
    }
    Esum += e;
    Wsum += w;
  }
  if (E) *E = 4. * eps * Esum / 2.;
  if (W) *W = Wsum / 2.;
}

void orc_forces_fp64(int N, const float* pos4, double L, int periodic,
                     double eps, double sigma, double skip_r2, double* f3,
                     double* ew, double* E, double* W, int nthreads)
{
  double* x = (double*)malloc((size_t)3 * N * sizeof(double));
  for (int i = 0; i < N; ++i)
    for (int c = 0; c < 3; ++c) x[3 * i + c] = pos4[4 * i + c];
  fp64_forces_core(N, x, L, periodic, eps, sigma, skip_r2, f3, ew, E, W, nthreads);
  free(x);
}

void orc_forces_fp64d(int N, const double* pos3, double L, int periodic,
                      double eps, double sigma, double* f3, double* ew,
                      double* E, double* W, int nthreads)
{
  fp64_forces_core(N, pos3, L, periodic, eps, sigma, 0., f3, ew, E, W, nthreads);
}

/* ---------------- RDF ---------------- */

void orc_rdf_edges2(int nbins, double rmax, float* edges2)
{
  double dr = rmax / nbins;
  for (int k = 0; k <= nbins; ++k) {
    double e = (double)k * dr;
    edges2[k] = (float)(e * e);
  }
}

/* MDSystem.cu:26-35 with nvcc's default contraction of the r^2 sum */
static inline float gpu_rule_r2(const float* pi, const float* pj, float Lf, int periodic)
{
  float r[3];
  for (int c = 0; c < 3; ++c) {
    float d = pi[c] - pj[c];
    if (periodic) {
      float q = d / Lf;
      long k = lroundf(q);
      float m = Lf * (float)k; /* exact for |k| <= 1 */
      d = d - m;
    }
    r[c] = d;
  }
  return fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0]));
}

static inline float cpu_rule_r2(const float* pi, const float* pj, double L, int periodic)
{
  float rx = pi[0] - pj[0], ry = pi[1] - pj[1], rz = pi[2] - pj[2];
  if (periodic) {
    rx = ref_minimage(rx, L);
    ry = ref_minimage(ry, L);
    rz = ref_minimage(rz, L);
  }
  return rx * rx + ry * ry + rz * rz;
}

void orc_rdf(int N, const float* pos, double L, int periodic, int rule,
             int nbins, float dr2, double rmax, int64_t* counts, int nthreads)
{
  float* edges2 = NULL;
  if (rule == ORC_RDF_RBINS) {
    edges2 = (float*)malloc((size_t)(nbins + 1) * sizeof(float));
    orc_rdf_edges2(nbins, rmax, edges2);
  }
  float Lf = (float)L;
  if (nthreads <= 0) {
#ifdef _OPENMP
    nthreads = omp_get_num_procs();
#else
    nthreads = 1;
#endif
  }
#pragma omp parallel num_threads(nthreads)
  {
    int64_t* loc = (int64_t*)calloc((size_t)nbins, sizeof(int64_t));
#pragma omp for schedule(dynamic, 16)
    for (int i = 0; i < N; ++i) {
      for (int j = 0; j < N; ++j) {
        const float* pi = pos + 4 * i;
        const float* pj = pos + 4 * j;
        if (rule == ORC_RDF_REF_GPU) {
          float r2 = gpu_rule_r2(pi, pj, Lf, periodic);
          if (r2 > 1e-6f) {
            float q = r2 / dr2;
            int b = (int)floorf(q);
            if (b > nbins - 1) b = nbins - 1;
            loc[b]++;
          }
          continue;
        }
        if (j == i) continue;
        float r2 = cpu_rule_r2(pi, pj, L, periodic);
        if (rule == ORC_RDF_REF_CPU) {
          int b = (int)floor((double)r2 / (double)dr2);
          if (b >= 0 && b < nbins) loc[b]++;
        } else {
          if (!(r2 < edges2[nbins])) continue;
          int lo = 0, hi = nbins; /* edges2[lo] <= r2 < edges2[hi] */
          while (hi - lo > 1) {
            int mid = (lo + hi) / 2;
            if (edges2[mid] <= r2) lo = mid;
            else hi = mid;
          }
          loc[lo]++;
        }
      }
    }
#pragma omp critical
    for (int b = 0; b < nbins; ++b) counts[b] += loc[b];
    free(loc);
  }
  free(edges2);
}

/* ---------------- full system ---------------- */

int orc_sys_init(orc_sys* s, int N, double T0, double rho, int canonical, int bc)
{
  memset(s, 0, sizeof(*s));
  if (N <= 0 || !(rho > 0.)) return -1;
  s->N = N;
  s->T0 = T0;
  s->rho = rho;
  s->canonical = canonical;
  s->bc = bc;
  s->L = orc_box_length(N, rho); /* :70 */
  s->T = T0;                     /* :72 */
  s->rdf_dr2 = orc_rdf_dr2(N);   /* :93-95 */
  s->pos = (float*)calloc((size_t)4 * N, sizeof(float));
  s->vel = (float*)calloc((size_t)4 * N, sizeof(float));
  s->frc = (float*)calloc((size_t)4 * N, sizeof(float));
  s->nthreads = 1;
  return (s->pos && s->vel && s->frc) ? 0 : -2;
}

void orc_sys_free(orc_sys* s)
{
  free(s->pos);
  free(s->vel);
  free(s->frc);
  s->pos = s->vel = s->frc = NULL;
}

void orc_sys_forces(orc_sys* s) /* :232-311 */
{
  if (s->nthreads > 1)
    orc_forces_ref_omp(s->N, s->pos, s->L, s->bc, s->rdf_dr2, s->frc, &s->V,
                       &s->P, &s->Pshear, s->rdf, s->nthreads);
  else
    orc_forces_ref(s->N, s->pos, s->L, s->bc, s->rdf_dr2, s->frc, &s->V, &s->P,
                   &s->Pshear, s->rdf);
}

void orc_sys_params(orc_sys* s) /* :325-359, CPU branch */
{
  const int N = s->N;
  const float* v = s->vel;
  s->K = 0.;
  for (int i = 0; i < 4 * N; i += 4) {
    float k2 = v[i] * v[i] + v[i + 1] * v[i + 1] + v[i + 2] * v[i + 2];
    s->K += k2 / 2.;
    float ps = -v[i] * v[i + 1];
    s->Pshear += ps;
  }
  s->T = 2. * s->K / 3. / N;
  s->P += N * s->T;
  s->P /= (N / s->rho);
  s->U = s->K + s->V;
  s->Pshear /= (N / s->rho);
  s->av_iters++;
  s->av_U_tot += s->U;
  s->av_p_tot += s->P;
  s->av_T_tot += s->T;
}

void orc_sys_apply_bc(orc_sys* s) /* :406-436 */
{
  const int N = s->N;
  float* p = s->pos;
  float* v = s->vel;
  const double L = s->L;
  if (s->bc == 2) return;
  if (s->bc == 0) {
    for (int i = 0; i < 4 * N; i += 4)
      for (int c = 0; c < 3; ++c) {
        if (p[i + c] < 0.) p[i + c] = (float)(p[i + c] + L);
        if (p[i + c] > L) p[i + c] = (float)(p[i + c] - L);
      }
  } else {
    for (int i = 0; i < 4 * N; i += 4)
      for (int c = 0; c < 3; ++c) {
        if (p[i + c] < 0. && v[i + c] < 0) v[i + c] = -v[i + c];
        if (p[i + c] > L && v[i + c] > 0) v[i + c] = -v[i + c];
      }
  }
}

void orc_sys_integrate(orc_sys* s, double dt) /* :438-510, :578-582 */
{
  const int N = s->N;
  float* p = s->pos;
  float* v = s->vel;
  float* f = s->frc;
  if (!s->canonical) {
    for (int i = 0; i < 4 * N; i += 4)
      for (int c = 0; c < 3; ++c) {
        float fc = f[i + c];
        p[i + c] = (float)(p[i + c] + (dt * v[i + c] + dt * dt * fc / 2.)); /* :447-449 */
        v[i + c] = (float)(v[i + c] + dt * fc / 2.);                        /* :451-453 */
      }
    orc_sys_forces(s);
    for (int i = 0; i < 4 * N; i += 4)
      for (int c = 0; c < 3; ++c) v[i + c] = (float)(v[i + c] + dt * f[i + c] / 2.); /* :460-462 */
  } else {
    float* tF = (float*)malloc((size_t)4 * N * sizeof(float));
    float* tV = (float*)calloc((size_t)4 * N, sizeof(float));
    for (int i = 0; i < 4 * N; i += 4)
      for (int c = 0; c < 3; ++c) {
        float fc = f[i + c];
        p[i + c] = (float)(p[i + c] + (dt * v[i + c] + dt * dt * fc / 2.)); /* :473-475 */
        tF[i + c] = 0.5f * f[i + c];                                        /* :477-479 */
      }
    orc_sys_forces(s);
    for (int i = 0; i < 4 * N; i += 4)
      for (int c = 0; c < 3; ++c) tF[i + c] += 0.5f * f[i + c];             /* :486-488 */
    for (int i = 0; i < 4 * N; i += 4)
      for (int c = 0; c < 3; ++c) tV[i + c] = (float)(v[i + c] + dt * tF[i + c] / 2.); /* :493-495 */
    double Tkin = orc_kinetic_temperature(N, tV);                            /* :498 */
    double chi = sqrt(s->T0 / Tkin);                                         /* :499 */
    for (int i = 0; i < 4 * N; i += 4)
      for (int c = 0; c < 3; ++c)
        v[i + c] = (float)((2. * chi - 1.) * v[i + c] + chi * dt * tF[i + c]); /* :503-505 */
    free(tF);
    free(tV);
  }
  orc_sys_apply_bc(s); /* :578 */
  orc_sys_params(s);   /* :579 */
  s->t += dt;          /* :582 */
}

/* ---------------- velocity distribution ---------------- */

/* Speed histogram of MDSystem::initvelo / updatevelo (MDSystem.cpp:651-694):
 * bin = (int)(sqrt((double)fl32(vx*vx + vy*vy + vz*vz)) / step), counted
 * when below nbins (the reference's maxind).  sqrt of the float sum resolves
 * to the double sqrt there (<cmath>, no using-directive in MDSystem.cpp). */
void orc_velocity_histogram(int N, const float* vel4, double step, int nbins, int64_t* counts)
{
  for (int b = 0; b < nbins; ++b) counts[b] = 0;
  for (int i = 0; i < 4 * N; i += 4) {
    const float s = vel4[i] * vel4[i] + vel4[i + 1] * vel4[i + 1] + vel4[i + 2] * vel4[i + 2];
    const int t = (int)(sqrt((double)s) / step);
    if (t < nbins) counts[t]++;
  }
}

/* ---------------- FP64 step for the precision study ---------------- */

/* Integrate(dt) of MDSystem.cpp:442-510 with every quantity in double:
 * EVN :442-464, TVN :465-510 (KineticTemperature :361-373 over the
 * half-step velocities), then the boundary conditions :406-436 applied
 * after the step as in :578. */
void orc_fp64_step(int N, double* x, double* v, double* f, double L, int bc, int canonical,
                   double T0, double eps, double sigma, double dt, double* obs3, int nthreads)
{
  double* tf = NULL;
  for (int i = 0; i < 3 * N; ++i) x[i] += dt * v[i] + dt * dt * f[i] / 2.;
  if (!canonical) {
    for (int i = 0; i < 3 * N; ++i) v[i] += dt * f[i] / 2.;
  } else {
    tf = (double*)malloc(sizeof(double) * 3 * (size_t)N);
    for (int i = 0; i < 3 * N; ++i) tf[i] = 0.5 * f[i];
  }
  double E = 0., W = 0.;
  orc_forces_fp64d(N, x, L, bc == 0, eps, sigma, f, NULL, &E, &W, nthreads);
  if (!canonical) {
    for (int i = 0; i < 3 * N; ++i) v[i] += dt * f[i] / 2.;
  } else {
    double tv2 = 0.;
    for (int i = 0; i < 3 * N; ++i) {
      tf[i] += 0.5 * f[i];
      const double tv = v[i] + dt * tf[i] / 2.;
      tv2 += tv * tv;
    }
    const double chi = sqrt(T0 / (tv2 / 3. / N));
    for (int i = 0; i < 3 * N; ++i) v[i] = (2. * chi - 1.) * v[i] + chi * dt * tf[i];
    free(tf);
  }
  if (bc == 0) {
    for (int i = 0; i < 3 * N; ++i) {
      if (x[i] < 0.) x[i] += L;
      if (x[i] > L) x[i] -= L;
    }
  } else if (bc == 1) {
    for (int i = 0; i < 3 * N; ++i) {
      if (x[i] < 0. && v[i] < 0.) v[i] = -v[i];
      if (x[i] > L && v[i] > 0.) v[i] = -v[i];
    }
  }
  double K = 0.;
  for (int i = 0; i < 3 * N; ++i) K += v[i] * v[i];
  obs3[0] = E;
  obs3[1] = K / 2.;
  obs3[2] = W;
}

/* ---------------- adjoint (FP64) ----------------
 * Pair quantities for the LJ pair of MDSystem.cpp:289-298 with eps, sigma:
 *   phi(r^2)  = 24 eps (2 s^12 r^-14 - s^6 r^-8)        F_i = sum_j phi r_ij
 *   phi'      = d phi / d(r^2) = 24 eps (-14 s^12 r^-16 + 4 s^6 r^-10)
 *   psi(r^2)  = 24 eps (2 s^12 r^-12 - s^6 r^-6)        W = sum_{i<j} psi
 *   psi'      = 24 eps (-12 s^12 r^-14 + 3 s^6 r^-8)
 *   (H z)_i   = sum_j [phi (z_i - z_j) + 2 phi' (r.(z_i - z_j)) r]
 *   dW/dx_i   = sum_j 2 psi' r_ij
 *   dE/ds     = sum_{i<j} 4 eps (12 s^11 r^-12 - 6 s^5 r^-6)
 *   dW/ds     = sum_{i<j} 24 eps (24 s^11 r^-12 - 6 s^5 r^-6)
 *   dphi/ds   = 24 eps (24 s^11 r^-14 - 6 s^5 r^-8)
 */
void orc_adjoint_pass_fp64(int N, const double* x, const double* z, double L, int periodic,
                           double eps, double s, double* hz3, double* f3, double* gw3,
                           double* scal8, int nthreads)
{
  double E = 0., W = 0., Es = 0., Ws = 0., Fsz = 0., Fz = 0.;
  const double s6 = pow(s, 6.), s12 = s6 * s6, s5 = s6 / s, s11 = s12 / s;
  if (nthreads <= 0) {
#ifdef _OPENMP
    nthreads = omp_get_num_procs();
#else
    nthreads = 1;
#endif
  }
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads) reduction(+ : E, W, Es, Ws, Fsz, Fz)
  for (int i = 0; i < N; ++i) {
    double h[3] = {0., 0., 0.}, f[3] = {0., 0., 0.}, g[3] = {0., 0., 0.};
    for (int j = 0; j < N; ++j) {
      if (j == i) continue;
      double r[3], dz[3];
      for (int c = 0; c < 3; ++c) {
        r[c] = x[3 * i + c] - x[3 * j + c];
        if (periodic) r[c] -= L * round(r[c] / L);
        dz[c] = z[3 * i + c] - z[3 * j + c];
      }
      const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      const double i2 = 1. / r2, i4 = i2 * i2, i6 = i4 * i2, i8 = i4 * i4, i10 = i8 * i2;
      const double i12 = i6 * i6, i14 = i12 * i2, i16 = i8 * i8;
      const double phi = 24. * eps * (2. * s12 * i14 - s6 * i8);
      const double dphi = 24. * eps * (-14. * s12 * i16 + 4. * s6 * i10);
      const double dpsi = 24. * eps * (-12. * s12 * i14 + 3. * s6 * i8);
      const double phis = 24. * eps * (24. * s11 * i14 - 6. * s5 * i8);
      const double rdz = r[0] * dz[0] + r[1] * dz[1] + r[2] * dz[2];
      for (int c = 0; c < 3; ++c) {
        h[c] += phi * dz[c] + 2. * dphi * rdz * r[c];
        f[c] += phi * r[c];
        g[c] += 2. * dpsi * r[c];
      }
      /* per ordered pair: half of each unordered-pair quantity */
      E += 0.5 * 4. * eps * (s12 * i12 - s6 * i6);
      W += 0.5 * 24. * eps * (2. * s12 * i12 - s6 * i6);
      Es += 0.5 * 4. * eps * (12. * s11 * i12 - 6. * s5 * i6);
      Ws += 0.5 * 24. * eps * (24. * s11 * i12 - 6. * s5 * i6);
      Fsz += 0.5 * phis * rdz;
    }
    for (int c = 0; c < 3; ++c) {
      if (hz3) hz3[3 * i + c] = h[c];
      if (f3) f3[3 * i + c] = f[c];
      if (gw3) gw3[3 * i + c] = g[c];
      Fz += f[c] * z[3 * i + c];
    }
  }
  if (scal8) {
    scal8[0] = E;
    scal8[1] = W;
    scal8[2] = Es;
    scal8[3] = Ws;
    scal8[4] = Fsz;
    scal8[5] = Fz;
    scal8[6] = scal8[7] = 0.;
  }
}

/* states x_m, v_m for m = 0..M (double[(M+1)*3N] each); E_m, W_m for m = 0..M */
static void fp64_traj(int N, const double* x0, const double* v0, double L, int periodic,
                      double eps, double s, double dt, int M, double* xs, double* vs,
                      double* Es, double* Ws, int nthreads)
{
  const int n3 = 3 * N;
  double* f = (double*)malloc(sizeof(double) * n3);
  memcpy(xs, x0, sizeof(double) * n3);
  memcpy(vs, v0, sizeof(double) * n3);
  orc_forces_fp64d(N, xs, L, periodic, eps, s, f, NULL, &Es[0], &Ws[0], nthreads);
  for (int m = 1; m <= M; ++m) {
    const double* xp = xs + (size_t)(m - 1) * n3;
    const double* vp = vs + (size_t)(m - 1) * n3;
    double* xn = xs + (size_t)m * n3;
    double* vn = vs + (size_t)m * n3;
    for (int k = 0; k < n3; ++k) {
      xn[k] = xp[k] + dt * vp[k] + 0.5 * dt * dt * f[k];
      vn[k] = vp[k] + 0.5 * dt * f[k];
    }
    orc_forces_fp64d(N, xn, L, periodic, eps, s, f, NULL, &Es[m], &Ws[m], nthreads);
    for (int k = 0; k < n3; ++k) vn[k] += 0.5 * dt * f[k];
  }
  free(f);
}

static void traj_loss_from(int N, int M, double L, const double* vs, const double* Es,
                           const double* Ws, double pt, double wp, double* out3)
{
  const int n3 = 3 * N;
  const double vol = L * L * L;
  double se = 0., sp = 0.;
  for (int m = 1; m <= M; ++m) {
    double v2 = 0.;
    const double* v = vs + (size_t)m * n3;
    for (int k = 0; k < n3; ++k) v2 += v[k] * v[k];
    se += Es[m];
    sp += (Ws[m] / 3. + v2 / 3.) / vol;
  }
  const double mE = se / M, mP = sp / M;
  out3[0] = mE + wp * (mP - pt) * (mP - pt);
  out3[1] = mE;
  out3[2] = mP;
}

void orc_traj_loss_fp64(int N, const double* x0, const double* v0, double L, int periodic,
                        double eps, double s, double dt, int M, double pt, double wp,
                        double* out3, int nthreads)
{
  const size_t n3 = (size_t)3 * N;
  double* xs = (double*)malloc(sizeof(double) * n3 * (M + 1));
  double* vs = (double*)malloc(sizeof(double) * n3 * (M + 1));
  double* Es = (double*)malloc(sizeof(double) * (M + 1));
  double* Ws = (double*)malloc(sizeof(double) * (M + 1));
  fp64_traj(N, x0, v0, L, periodic, eps, s, dt, M, xs, vs, Es, Ws, nthreads);
  traj_loss_from(N, M, L, vs, Es, Ws, pt, wp, out3);
  free(xs);
  free(vs);
  free(Es);
  free(Ws);
}

void orc_traj_grad_fp64(int N, const double* x0, const double* v0, double L, int periodic,
                        double eps, double s, double dt, int M, double pt, double wp,
                        double* out3, double* grad2, double* dx0, int nthreads)
{
  const size_t n3 = (size_t)3 * N;
  double* xs = (double*)malloc(sizeof(double) * n3 * (M + 1));
  double* vs = (double*)malloc(sizeof(double) * n3 * (M + 1));
  double* Es = (double*)malloc(sizeof(double) * (M + 1));
  double* Ws = (double*)malloc(sizeof(double) * (M + 1));
  fp64_traj(N, x0, v0, L, periodic, eps, s, dt, M, xs, vs, Es, Ws, nthreads);
  traj_loss_from(N, M, L, vs, Es, Ws, pt, wp, out3);
  const double vol = L * L * L, invM = 1. / M, cP = 2. * wp * (out3[2] - pt);
  const double cw = cP / (3. * vol), gvs = invM * cP * (2. / 3.) / vol;
  double* u = (double*)calloc(n3, sizeof(double));
  double* lv = (double*)malloc(sizeof(double) * n3);
  double* z = (double*)malloc(sizeof(double) * n3);
  double* w = (double*)malloc(sizeof(double) * n3);
  double* hz = (double*)malloc(sizeof(double) * n3);
  double* f = (double*)malloc(sizeof(double) * n3);
  double* gw = (double*)malloc(sizeof(double) * n3);
  double sc[8];
  double ge = 0., gs = 0.;
  /* top: lv_M = g_v(M), z_M = dt/2 lv_M */
  for (size_t k = 0; k < n3; ++k) {
    lv[k] = gvs * vs[(size_t)M * n3 + k];
    z[k] = 0.5 * dt * lv[k];
  }
  for (int m = M; m >= 0; --m) {
    orc_adjoint_pass_fp64(N, xs + (size_t)m * n3, z, L, periodic, eps, s, hz, f, gw, sc, nthreads);
    ge += sc[5] / eps;
    gs += sc[4];
    if (m >= 1) {
      ge += invM * (sc[0] / eps + cw * sc[1] / eps);
      gs += invM * (sc[2] + cw * sc[3]);
      for (size_t k = 0; k < n3; ++k) {
        u[k] += invM * (-f[k] + cw * gw[k]) + hz[k];            /* u_{m-1} */
        w[k] = 0.5 * dt * dt * u[k] + 0.5 * dt * lv[k];         /* w_{m-1} */
        const double gv = (m - 1 >= 1) ? gvs * vs[(size_t)(m - 1) * n3 + k] : 0.;
        lv[k] = dt * u[k] + lv[k] + gv;                         /* lv_{m-1} */
        z[k] = (m - 1 >= 1) ? w[k] + 0.5 * dt * lv[k] : w[k];
      }
    } else if (dx0) {
      for (size_t k = 0; k < n3; ++k) dx0[k] = u[k] + hz[k];
    }
  }
  grad2[0] = ge;
  grad2[1] = gs;
  free(u);
  free(lv);
  free(z);
  free(w);
  free(hz);
  free(f);
  free(gw);
  free(xs);
  free(vs);
  free(Es);
  free(Ws);
}
