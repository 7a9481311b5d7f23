/*
 * fcm_oracle.c -- FP64 CPU restatement of the FCM mobility solve.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows FCMfunctions.cpp:
 *   wavenumbers            FCMfunctions.cpp:129-162
 *   GA_setup               FCMfunctions.cpp:536-547
 *   gaussian_setup         FCMfunctions.cpp:549-593
 *   force distribution     FCMfunctions.cpp:596-646
 *   flow_solve             FCMfunctions.cpp:648-736
 *   particle_velocities    FCMfunctions.cpp:740-770
 *   particle_rotations     FCMfunctions.cpp:774-808
 *   friction_mobility      FCMfunctions.cpp:865-880
 * and replaces FFTW 2.1.5 rfftwnd (FCMfunctions.cpp:360-406) by a
 * mixed-radix Cooley-Tukey restated from the DFT definition.
 *
 * Parity pins (tests/test_oracle_*.py): or_fft1d / or_rfft3d / or_irfft3d
 * against numpy.fft fixtures and a direct O(N^2) DFT, Parseval and
 * inverse(forward) = N*identity; or_fcm_project against the closed-form
 * sinusoidal body force the reference's own dormant test uses
 * (FCMfunctions.cpp:252-286) and k.u = 0; spread/interp through adjointness
 * <F1, M F2> = <F2, M F1>, linearity, whole-cell translation invariance,
 * grid-sum normalisation, brute-force per-point spread on a tiny grid, and
 * the single-segment Hasimoto self-mobility.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846264338327950288

static double g_times[5];

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void or_fcm_last_times(double *t5) { memcpy(t5, g_times, sizeof g_times); }

/* threads the OpenMP loops use (1 without OpenMP) */
int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed radix, out-of-place, twiddles from a table.    */
/* ------------------------------------------------------------------ */
static int smallest_factor(int n) {
  if (n % 4 == 0 && n > 4) return 4;
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

/* out[k] = sum_j in[j*stride] w^(j k), w = exp(sign 2 pi i / n); tw holds
 * exp(sign 2 pi i m / ntop) for m < ntop, and n divides ntop. */
static void fft_rec(int n, int stride, const double *in, double *out,
                    const double *tw, int ntop) {
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  int p = smallest_factor(n);
  int m = n / p;
  /* sub-transforms of the p decimated sequences */
  for (int r = 0; r < p; ++r)
    fft_rec(m, stride * p, in + 2 * r * stride, out + 2 * r * m, tw, ntop);
  /* combine: X[k + q m] = sum_r w_n^{r(k+qm)} Y_r[k] */
  double tmp[2 * 64];
  double *t = tmp;
  double *heap = NULL;
  if (p > 64) {
    heap = (double *)malloc(sizeof(double) * 2 * p);
    t = heap;
  }
  int step = ntop / n;
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) {
      t[2 * r] = out[2 * (r * m + k)];
      t[2 * r + 1] = out[2 * (r * m + k) + 1];
    }
    for (int q = 0; q < p; ++q) {
      double sr = 0.0, si = 0.0;
      int kk = k + q * m;
      for (int r = 0; r < p; ++r) {
        long e = ((long)r * kk) % n;
        double wr = tw[2 * (e * step)], wi = tw[2 * (e * step) + 1];
        sr += t[2 * r] * wr - t[2 * r + 1] * wi;
        si += t[2 * r] * wi + t[2 * r + 1] * wr;
      }
      out[2 * kk] = sr;
      out[2 * kk + 1] = si;
    }
  }
  free(heap);
}

typedef struct {
  int n;
  int sign;
  double *tw;
} fft_plan;

static void plan_init(fft_plan *pl, int n, int sign) {
  pl->n = n;
  pl->sign = sign;
  pl->tw = (double *)malloc(sizeof(double) * 2 * n);
  for (int m = 0; m < n; ++m) {
    /* angle reduced exactly in integers before the trig call */
    double ang = 2.0 * OR_PI * (double)m / (double)n;
    pl->tw[2 * m] = cos(ang);
    pl->tw[2 * m + 1] = sign * sin(ang);
  }
}

static void plan_free(fft_plan *pl) { free(pl->tw); }

static void plan_exec(const fft_plan *pl, double *data, double *scratch) {
  fft_rec(pl->n, 1, data, scratch, pl->tw, pl->n);
  memcpy(data, scratch, sizeof(double) * 2 * pl->n);
}

void or_fft1d(int n, double *data, int sign) {
  fft_plan pl;
  plan_init(&pl, n, sign);
  double *s = (double *)malloc(sizeof(double) * 2 * n);
  plan_exec(&pl, data, s);
  free(s);
  plan_free(&pl);
}

/* Transform along x and y of a half spectrum c[nx][ny][nzh]. */
static void fft_xy(int nx, int ny, int nzh, double *c, int sign) {
  int nmax = nx > ny ? nx : ny;
  fft_plan py, px;
  plan_init(&py, ny, sign);
  plan_init(&px, nx, sign);
  /* each x plane [y][k] (and each y plane [x][k]) is copied into nzh
   * contiguous lines, transformed line by line and copied back: the same
   * per-line transforms as a strided walk, with cache-friendly access */
#pragma omp parallel
  {
    double *blk = (double *)malloc(sizeof(double) * 2 * (size_t)nmax * nzh);
    double *s = (double *)malloc(sizeof(double) * 2 * nmax);
#pragma omp for schedule(static)
    for (int x = 0; x < nx; ++x) {
      double *P = c + 2 * (size_t)x * ny * nzh;
      for (int y = 0; y < ny; ++y)
        for (int k = 0; k < nzh; ++k) {
          blk[2 * ((size_t)k * ny + y)] = P[2 * ((size_t)y * nzh + k)];
          blk[2 * ((size_t)k * ny + y) + 1] = P[2 * ((size_t)y * nzh + k) + 1];
        }
      for (int k = 0; k < nzh; ++k) plan_exec(&py, blk + 2 * (size_t)k * ny, s);
      for (int y = 0; y < ny; ++y)
        for (int k = 0; k < nzh; ++k) {
          P[2 * ((size_t)y * nzh + k)] = blk[2 * ((size_t)k * ny + y)];
          P[2 * ((size_t)y * nzh + k) + 1] = blk[2 * ((size_t)k * ny + y) + 1];
        }
    }
#pragma omp for schedule(static)
    for (int y = 0; y < ny; ++y) {
      for (int x = 0; x < nx; ++x) {
        const double *P = c + 2 * ((size_t)x * ny + y) * nzh;
        for (int k = 0; k < nzh; ++k) {
          blk[2 * ((size_t)k * nx + x)] = P[2 * k];
          blk[2 * ((size_t)k * nx + x) + 1] = P[2 * k + 1];
        }
      }
      for (int k = 0; k < nzh; ++k) plan_exec(&px, blk + 2 * (size_t)k * nx, s);
      for (int x = 0; x < nx; ++x) {
        double *P = c + 2 * ((size_t)x * ny + y) * nzh;
        for (int k = 0; k < nzh; ++k) {
          P[2 * k] = blk[2 * ((size_t)k * nx + x)];
          P[2 * k + 1] = blk[2 * ((size_t)k * nx + x) + 1];
        }
      }
    }
    free(blk);
    free(s);
  }
  plan_free(&py);
  plan_free(&px);
}

/* The OpenMP loops below split independent lines, planes or segments
 * between threads; every output value is computed by one thread in the
 * serial order, so results do not depend on the thread count. */
void or_rfft3d(int nx, int ny, int nz, const double *r, double *c) {
  int nzh = nz / 2 + 1;
  fft_plan pz;
  plan_init(&pz, nz, -1);
#pragma omp parallel
  {
    double *line = (double *)malloc(sizeof(double) * 2 * nz);
    double *s = (double *)malloc(sizeof(double) * 2 * nz);
#pragma omp for schedule(static)
    for (long row = 0; row < (long)nx * ny; ++row) {
      for (int z = 0; z < nz; ++z) {
        line[2 * z] = r[row * nz + z];
        line[2 * z + 1] = 0.0;
      }
      plan_exec(&pz, line, s);
      memcpy(c + 2 * row * nzh, line, sizeof(double) * 2 * nzh);
    }
    free(line);
    free(s);
  }
  plan_free(&pz);
  fft_xy(nx, ny, nzh, c, -1);
}

void or_irfft3d(int nx, int ny, int nz, const double *cin, double *r) {
  int nzh = nz / 2 + 1;
  size_t csz = (size_t)nx * ny * nzh * 2;
  double *c = (double *)malloc(sizeof(double) * csz);
  memcpy(c, cin, sizeof(double) * csz);
  fft_xy(nx, ny, nzh, c, +1);
  fft_plan pz;
  plan_init(&pz, nz, +1);
#pragma omp parallel
  {
    double *line = (double *)malloc(sizeof(double) * 2 * nz);
    double *s = (double *)malloc(sizeof(double) * 2 * nz);
#pragma omp for schedule(static)
    for (long row = 0; row < (long)nx * ny; ++row) {
      const double *h = c + 2 * row * nzh;
      for (int k = 0; k < nzh; ++k) {
        line[2 * k] = h[2 * k];
        line[2 * k + 1] = h[2 * k + 1];
      }
      /* a real 1D signal's DC and Nyquist coefficients are real */
      line[1] = 0.0;
      if (nz % 2 == 0) line[2 * (nz / 2) + 1] = 0.0;
      for (int k = nzh; k < nz; ++k) { /* Hermitian completion */
        line[2 * k] = h[2 * (nz - k)];
        line[2 * k + 1] = -h[2 * (nz - k) + 1];
      }
      plan_exec(&pz, line, s);
      for (int z = 0; z < nz; ++z) r[row * nz + z] = line[2 * z];
    }
    free(line);
    free(s);
  }
  plan_free(&pz);
  free(c);
}

/* ------------------------------------------------------------------ */
/* FCM                                                                  */
/* ------------------------------------------------------------------ */
typedef struct {
  int ngd, ngdh;
  double anorm, anorm2, anormdip, anormdip2, sigmadipsq;
} gconst;

/* multi_filament_header.hpp:50-60 */
static gconst make_gconst(const fs_fcm_params *p) {
  gconst g;
  g.ngd = p->ngd;
  g.ngdh = p->ngd / 2;
  g.sigmadipsq = p->sigma_dip * p->sigma_dip;
  g.anormdip = 1.0 / sqrt(2.0 * OR_PI * g.sigmadipsq);
  g.anormdip2 = 2.0 * g.sigmadipsq;
  double sigmasq = p->sigma * p->sigma;
  g.anorm = 1.0 / sqrt(2.0 * OR_PI * sigmasq);
  g.anorm2 = 2.0 * sigmasq;
  return g;
}

/* gaussian_setup for one segment and one axis (FCMfunctions.cpp:564-591):
 * stencil centre by truncation, wrapped indices, weights. */
static void axis_tables(const gconst *g, double y, double dx, int n, int *ind,
                        double *ga, double *gd, double *gr) {
  int c = (int)(y / dx);
  for (int i = 0; i < g->ngd; ++i) {
    int xg = c - g->ngdh + (i + 1);
    ind[i] = xg - n * (int)floor((double)xg / (double)n);
    double xx = ((double)xg) * dx - y;
    ga[i] = g->anorm * exp(-xx * xx / g->anorm2);
    gd[i] = g->anormdip * exp(-xx * xx / g->anormdip2);
    gr[i] = -xx / g->sigmadipsq;
  }
}

typedef struct {
  int *ind[3];
  double *ga[3], *gd[3], *gr[3];
} seg_tables;

static void tables_alloc(seg_tables *t, int ngd) {
  for (int d = 0; d < 3; ++d) {
    t->ind[d] = (int *)malloc(sizeof(int) * ngd);
    t->ga[d] = (double *)malloc(sizeof(double) * ngd);
    t->gd[d] = (double *)malloc(sizeof(double) * ngd);
    t->gr[d] = (double *)malloc(sizeof(double) * ngd);
  }
}

static void tables_free(seg_tables *t) {
  for (int d = 0; d < 3; ++d) {
    free(t->ind[d]);
    free(t->ga[d]);
    free(t->gd[d]);
    free(t->gr[d]);
  }
}

static void tables_fill(const fs_fcm_params *p, const gconst *g,
                        const double *y, seg_tables *t) {
  int n[3] = {p->nx, p->ny, p->nz};
  for (int d = 0; d < 3; ++d)
    axis_tables(g, y[d], p->dx, n[d], t->ind[d], t->ga[d], t->gd[d], t->gr[d]);
}

/* GA_setup + mono_and_dipole_force_distribution
 * (FCMfunctions.cpp:536-547, 596-646) with GS = 0 (FCMfunctions.cpp:200). */
void or_fcm_spread(const fs_fcm_params *p, int np, const double *Y,
                   const double *F, const double *T, double *f3) {
  gconst g = make_gconst(p);
  size_t N = (size_t)p->nx * p->ny * p->nz;
  memset(f3, 0, sizeof(double) * 3 * N);
  double *fx = f3, *fy = f3 + N, *fz = f3 + 2 * N;
  /* each thread owns the x planes [x0, x1) and adds every segment's terms
   * on them in segment order, as the serial loop does */
#pragma omp parallel
  {
#ifdef _OPENMP
  const int nth = omp_get_num_threads(), tid = omp_get_thread_num();
#else
  const int nth = 1, tid = 0;
#endif
  const int x0 = (int)((long)p->nx * tid / nth);
  const int x1 = (int)((long)p->nx * (tid + 1) / nth);
  seg_tables t;
  tables_alloc(&t, g.ngd);
  for (int s = 0; s < np; ++s) {
    axis_tables(&g, Y[3 * s], p->dx, p->nx, t.ind[0], t.ga[0], t.gd[0],
                t.gr[0]);
    int hit = 0;
    for (int i = 0; i < g.ngd; ++i)
      hit |= t.ind[0][i] >= x0 && t.ind[0][i] < x1;
    if (!hit) continue;
    tables_fill(p, &g, Y + 3 * s, &t);
    double GA3 = 0.5 * T[3 * s + 2], GA4 = -0.5 * T[3 * s + 1],
           GA5 = 0.5 * T[3 * s + 0];
    double g11 = 0.0, g22 = 0.0, g33 = 0.0;
    double g12 = GA3, g21 = -GA3, g13 = GA4, g31 = -GA4, g23 = GA5,
           g32 = -GA5;
    for (int i = 0; i < g.ngd; ++i) {
      int ii = t.ind[0][i];
      if (ii < x0 || ii >= x1) continue;
      double xx = t.gr[0][i];
      for (int j = 0; j < g.ngd; ++j) {
        int jj = t.ind[1][j];
        double yy = t.gr[1][j];
        for (int k = 0; k < g.ngd; ++k) {
          int kk = t.ind[2][k];
          double zz = t.gr[2][k];
          double temp = t.ga[0][i] * t.ga[1][j] * t.ga[2][k];
          double temp2 = t.gd[0][i] * t.gd[1][j] * t.gd[2][k];
          size_t o = ((size_t)ii * p->ny + jj) * p->nz + kk;
          fx[o] = fx[o] + F[3 * s + 0] * temp +
                  (g11 * xx + g12 * yy + g13 * zz) * temp2;
          fy[o] = fy[o] + F[3 * s + 1] * temp +
                  (g21 * xx + g22 * yy + g23 * zz) * temp2;
          fz[o] = fz[o] + F[3 * s + 2] * temp +
                  (g31 * xx + g32 * yy + g33 * zz) * temp2;
        }
      }
    }
  }
  tables_free(&t);
  }
}

/* Wavenumbers (FCMfunctions.cpp:129-162): Nyquist index gets +N/2. */
static double wavenumber(int i, int n, double L) {
  double fac = 2.0 * OR_PI / L;
  return (i <= n / 2) ? i * fac : (i - n) * fac;
}

/* flow_solve (FCMfunctions.cpp:648-736) on [nx][ny][nzh] half spectra. The
 * reference omits 1/mu (exact only for mu = 1); it is applied here. */
void or_fcm_project(const fs_fcm_params *p, double *c3) {
  int nx = p->nx, ny = p->ny, nz = p->nz, nzh = nz / 2 + 1;
  size_t NC = (size_t)nx * ny * nzh;
  double *cx = c3, *cy = c3 + 2 * NC, *cz = c3 + 4 * NC;
  double totalpts = (double)nx * ny * nz;
  double smallx = 1e-18;
  double Lx = nx * p->dx, Ly = ny * p->dx, Lz = nz * p->dx;
  /* k = 0: zero the mean force (FCMfunctions.cpp:659-666) */
  cx[0] = cx[1] = cy[0] = cy[1] = cz[0] = cz[1] = 0.0;
#pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; ++x) {
    double q1 = wavenumber(x, nx, Lx);
    for (int y = 0; y < ny; ++y) {
      double q2 = wavenumber(y, ny, Ly);
      double qq = q2 * q2 + q1 * q1 + smallx;
      for (int k = 0; k < nzh; ++k) {
        double q3 = k * (2.0 * OR_PI / Lz);
        double norm = 1.0 / (qq + q3 * q3);
        size_t o = 2 * (((size_t)x * ny + y) * nzh + k);
        for (int part = 0; part < 2; ++part) {
          double f1 = cx[o + part], f2 = cy[o + part], f3 = cz[o + part];
          double kdotf = (q1 * f1 + q2 * f2 + q3 * f3) * norm;
          cx[o + part] = norm * (f1 - q1 * kdotf) / totalpts / p->mu;
          cy[o + part] = norm * (f2 - q2 * kdotf) / totalpts / p->mu;
          cz[o + part] = norm * (f3 - q3 * kdotf) / totalpts / p->mu;
        }
      }
    }
  }
  cx[0] = cx[1] = cy[0] = cy[1] = cz[0] = cz[1] = 0.0; /* :727-734 */
}

/* particle_velocities + particle_rotations (FCMfunctions.cpp:740-808). */
void or_fcm_interp(const fs_fcm_params *p, int np, const double *Y,
                   const double *u3, double *V, double *W) {
  gconst g = make_gconst(p);
  size_t N = (size_t)p->nx * p->ny * p->nz;
  const double *ux = u3, *uy = u3 + N, *uz = u3 + 2 * N;
  double norm = p->dx * p->dx * p->dx;
#pragma omp parallel
  {
  seg_tables t;
  tables_alloc(&t, g.ngd);
#pragma omp for schedule(static)
  for (int s = 0; s < np; ++s) {
    tables_fill(p, &g, Y + 3 * s, &t);
    double v0 = 0, v1 = 0, v2 = 0;
    for (int i = 0; i < g.ngd; ++i) {
      int ii = t.ind[0][i];
      for (int j = 0; j < g.ngd; ++j) {
        int jj = t.ind[1][j];
        for (int k = 0; k < g.ngd; ++k) {
          int kk = t.ind[2][k];
          double temp = t.ga[0][i] * t.ga[1][j] * t.ga[2][k] * norm;
          size_t o = ((size_t)ii * p->ny + jj) * p->nz + kk;
          v0 = v0 + ux[o] * temp;
          v1 = v1 + uy[o] * temp;
          v2 = v2 + uz[o] * temp;
        }
      }
    }
    double w0 = 0, w1 = 0, w2 = 0;
    for (int i = 0; i < g.ngd; ++i) {
      int ii = t.ind[0][i];
      double xx = t.gr[0][i];
      for (int j = 0; j < g.ngd; ++j) {
        int jj = t.ind[1][j];
        double yy = t.gr[1][j];
        for (int k = 0; k < g.ngd; ++k) {
          int kk = t.ind[2][k];
          double zz = t.gr[2][k];
          double temp = t.gd[0][i] * t.gd[1][j] * t.gd[2][k] * norm;
          size_t o = ((size_t)ii * p->ny + jj) * p->nz + kk;
          w0 = w0 - 0.5 * (uz[o] * yy - uy[o] * zz) * temp;
          w1 = w1 - 0.5 * (ux[o] * zz - uz[o] * xx) * temp;
          w2 = w2 - 0.5 * (uy[o] * xx - ux[o] * yy) * temp;
        }
      }
    }
    V[3 * s + 0] = v0;
    V[3 * s + 1] = v1;
    V[3 * s + 2] = v2;
    W[3 * s + 0] = w0;
    W[3 * s + 1] = w1;
    W[3 * s + 2] = w2;
  }
  tables_free(&t);
  }
}

void or_fcm_mobility(const fs_fcm_params *p, int np, const double *Y,
                     const double *F, const double *T, double *V, double *W) {
  size_t N = (size_t)p->nx * p->ny * p->nz;
  size_t NC = (size_t)p->nx * p->ny * (p->nz / 2 + 1);
  double *f3 = (double *)malloc(sizeof(double) * 3 * N);
  double *c3 = (double *)malloc(sizeof(double) * 6 * NC);
  double t0 = now_s();
  or_fcm_spread(p, np, Y, F, T, f3);
  double t1 = now_s();
  for (int c = 0; c < 3; ++c)
    or_rfft3d(p->nx, p->ny, p->nz, f3 + c * N, c3 + 2 * c * NC);
  double t2 = now_s();
  or_fcm_project(p, c3);
  double t3 = now_s();
  for (int c = 0; c < 3; ++c)
    or_irfft3d(p->nx, p->ny, p->nz, c3 + 2 * c * NC, f3 + c * N);
  double t4 = now_s();
  or_fcm_interp(p, np, Y, f3, V, W);
  double t5 = now_s();
  g_times[0] = t1 - t0;
  g_times[1] = t2 - t1;
  g_times[2] = t3 - t2;
  g_times[3] = t4 - t3;
  g_times[4] = t5 - t4;
  free(f3);
  free(c3);
}

/* Fcm::friction_mobility_solve (FCMfunctions.cpp:865-880). */
void or_fcm_friction(const fs_fcm_params *p, int np, const double *F,
                     const double *T, double *V, double *W) {
  double a1 = 1. / (6 * OR_PI * p->a * p->mu);
  double a2 = 1. / (8 * OR_PI * p->a * p->a * p->a * p->mu);
  for (int i = 0; i < 3 * np; ++i) {
    V[i] = a1 * F[i];
    W[i] = a2 * T[i];
  }
}
