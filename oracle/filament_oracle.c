/*
 * filament_oracle.c -- FP64 CPU restatement of the filament mechanics, the
 * BDF2 residual, the approximate Jacobian and the bad-Broyden timestep.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   quaternion algebra      UnitQuaternion.hpp:103-157, 227-242, 323-378,
 *                           411-421, 472-489
 *   bead kinematics         Bead.hpp:202-276
 *   filament forces         Filament.hpp:227-237, 271-402, 425-505
 *   residual + dexpinv      MobilityConstraints.cpp:109-143, 447-770
 *   Jacobian blocks         Filament.hpp:803-1120 (rcross, BlockC/D/E/F)
 *   Broyden solve/update    FilamentJacobianSolver.hpp:36-84,
 *                           main.cpp:494-687
 *   timestep                main.cpp:284-754
 *   assign_filament_data    FCMfunctions.cpp:882-918
 *   spring links            spring_link_functions.cpp:24-45, 131-138
 * arma::inv (Filament.hpp:862, LAPACK getrf/getri) is replaced by an LU
 * with partial pivoting restated from the textbook algorithm.
 *
 * Parity pins (tests/test_oracle_filament.py): quaternion identities
 * (|q| = 1, expm of an axis-angle equals the rotation-matrix action,
 * midpoint^2 = relative rotation, tangent = R(q) e_x); the elastic torque
 * of a link turned by theta about a body axis equals the closed form
 * R(q_m) diag(C, kap, kap) ((4/dL) sin(theta/4) e - p) (twist and both
 * bends, either sign, straight and turned past the preferred strain p) and
 * vanishes when every link carries its preferred strain; constraint forces
 * sum to zero; dexpinv inverts the exp-map derivative (finite differences);
 * the residual vanishes on exact friction-mobility translation; J * J^-1 =
 * I; the Broyden secant condition; the analytic Jacobian blocks against
 * finite differences of the residual: first-bead rows, BlockC, BlockD and
 * BlockF where the approximation is exact; BlockE exactly without elastic
 * moduli (constraint torques, bent and twisted filaments, nonzero lambda
 * and TAU), and on a straight untwisted filament with no preferred strain
 * its twist blocks are exact but for the sign of the left off-diagonal
 * block and its bend blocks exactly half the derivative (properties of the
 * reference's approximation, restated as it is); whole steps converge to
 * TOL; spring forces against Hooke's law on hand-placed beads, across the
 * periodic boundary, equal and opposite.
 * Parity unpinned: the preferred-strain terms of BlockE (beta's
 * -StrainTwist(2), the StrainTwist(0) rcross(n) + StrainTwist(1) rcross(b)
 * terms) and the curvature parts of beta and CrossPart on bent frames: they
 * are approximations with no exact counterpart to compare against, pinned
 * only by re-derivation from the source text and by Broyden convergence.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define OR_PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------ */
/* 3-vector and quaternion helpers (q = [scal, v0, v1, v2])            */
/* ------------------------------------------------------------------ */
static void cross3(const double *a, const double *b, double *c) {
  double c0 = a[1] * b[2] - a[2] * b[1];
  double c1 = a[2] * b[0] - a[0] * b[2];
  double c2 = a[0] * b[1] - a[1] * b[0];
  c[0] = c0;
  c[1] = c1;
  c[2] = c2;
}
static double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static double norm3(const double *a) { return sqrt(dot3(a, a)); }

/* p * this, in place on q (quaproFromLeftInPlace, UnitQuaternion.hpp:137) */
static void q_from_left(double *q, const double *p) {
  double s = q[0] * p[0] - dot3(q + 1, p + 1);
  double t[3];
  cross3(p + 1, q + 1, t);
  double v0 = q[0] * p[1] + p[0] * q[1] + t[0];
  double v1 = q[0] * p[2] + p[0] * q[2] + t[1];
  double v2 = q[0] * p[3] + p[0] * q[3] + t[2];
  q[0] = s;
  q[1] = v0;
  q[2] = v1;
  q[3] = v2;
}

/* (p^*) * this, in place (quaproFromLeftInPlaceConj, :146-153) */
static void q_from_left_conj(double *q, const double *p) {
  double s = q[0] * p[0] + dot3(q + 1, p + 1);
  double t[3];
  cross3(q + 1, p + 1, t);
  double v0 = -q[0] * p[1] + p[0] * q[1] + t[0];
  double v1 = -q[0] * p[2] + p[0] * q[2] + t[1];
  double v2 = -q[0] * p[3] + p[0] * q[3] + t[2];
  q[0] = s;
  q[1] = v0;
  q[2] = v1;
  q[3] = v2;
}

/* this * p, in place (quaproInPlace, :119-125) */
static void q_times(double *q, const double *p) {
  double s = q[0] * p[0] - dot3(q + 1, p + 1);
  double t[3];
  cross3(q + 1, p + 1, t);
  double v0 = q[0] * p[1] + p[0] * q[1] + t[0];
  double v1 = q[0] * p[2] + p[0] * q[2] + t[1];
  double v2 = q[0] * p[3] + p[0] * q[3] + t[2];
  q[0] = s;
  q[1] = v0;
  q[2] = v1;
  q[3] = v2;
}

/* expmInPlace (UnitQuaternion.hpp:227-242): q <- exp(v) q */
void or_quat_expm(double *q, const double *v) {
  double theta = norm3(v);
  if (theta > 1.e-12) {
    double tmp = sin(theta * .5) / theta;
    double qu[4] = {cos(theta * .5), tmp * v[0], tmp * v[1], tmp * v[2]};
    q_from_left(q, qu);
  }
}

/* assignMidPointOf (UnitQuaternion.hpp:472-489) with
 * quaSquareRootTimInPlace (:411-421). */
void or_quat_midpoint(const double *p1, const double *p2, double *out) {
  /* p2 * p1^* (quaproConj, :113-115) */
  double t[3];
  cross3(p2 + 1, p1 + 1, t);
  double s = p2[0] * p1[0] + dot3(p2 + 1, p1 + 1);
  double v0 = -p2[0] * p1[1] + p1[0] * p2[1] - t[0];
  double v1 = -p2[0] * p1[2] + p1[0] * p2[2] - t[1];
  double v2 = -p2[0] * p1[3] + p1[0] * p2[3] - t[2];
  if (s < -.9999999999999) {
    s = 0.;
    v0 = 0.;
    v1 = 0.;
    v2 = 1.;
  } else {
    s = sqrt(0.5 * (s + 1.));
    v0 /= (2. * s);
    v1 /= (2. * s);
    v2 /= (2. * s);
  }
  out[0] = s;
  out[1] = v0;
  out[2] = v1;
  out[3] = v2;
  q_times(out, p1);
}

/* assignTangentVec / NormalVec / BinormalVec (UnitQuaternion.hpp:341-378) */
void or_quat_tangent(const double *q, double *v) {
  double q22 = 2 * q[2] * q[2];
  double q33 = 2 * q[3] * q[3];
  v[0] = (1 - q22 - q33);
  v[1] = 2 * (q[1] * q[2] + q[0] * q[3]);
  v[2] = 2 * (q[1] * q[3] - q[0] * q[2]);
}
static void quat_normal(const double *q, double *v) {
  double q11 = 2 * q[1] * q[1];
  double q33 = 2 * q[3] * q[3];
  v[0] = 2 * (q[1] * q[2] - q[0] * q[3]);
  v[1] = (1 - q11 - q33);
  v[2] = 2 * (q[2] * q[3] + q[0] * q[1]);
}
static void quat_binormal(const double *q, double *v) {
  double q11 = 2 * q[1] * q[1];
  double q22 = 2 * q[2] * q[2];
  v[0] = 2 * (q[1] * q[3] + q[0] * q[2]);
  v[1] = 2 * (q[2] * q[3] - q[0] * q[1]);
  v[2] = (1 - q11 - q22);
}

/* RotateVector (UnitQuaternion.hpp:323-338) */
void or_quat_rotate(const double *q, double *v) {
  double vx = v[0], vy = v[1], vz = v[2];
  double q11 = 2 * q[1] * q[1];
  double q22 = 2 * q[2] * q[2];
  double q33 = 2 * q[3] * q[3];
  v[0] = (1 - q22 - q33) * vx + 2 * (q[1] * q[2] - q[0] * q[3]) * vy +
         2 * (q[1] * q[3] + q[0] * q[2]) * vz;
  v[1] = 2 * (q[1] * q[2] + q[0] * q[3]) * vx + (1 - q11 - q33) * vy +
         2 * (q[2] * q[3] - q[0] * q[1]) * vz;
  v[2] = 2 * (q[1] * q[3] - q[0] * q[2]) * vx +
         2 * (q[2] * q[3] + q[0] * q[1]) * vy + (1 - q11 - q22) * vz;
}

/* dexpinv (MobilityConstraints.cpp:109-143) */
void or_dexpinv(const double *u, const double *w, double *out) {
  double tmp = norm3(u);
  double uw[3];
  cross3(u, w, uw);
  if (tmp >= 1.e-12) {
    double tmp2 = tmp / sin(.5 * tmp);
    tmp2 *= cos(.5 * tmp);
    double c = .5 * tmp2 - 1.;
    double un[3] = {u[0] / tmp, u[1] / tmp, u[2] / tmp};
    double a[3], b[3];
    cross3(un, w, a);
    cross3(un, a, b);
    for (int d = 0; d < 3; ++d) out[d] = (w[d] - .5 * uw[d]) - c * b[d];
  } else {
    double c = 1. / 12;
    double b[3];
    cross3(u, uw, b);
    for (int d = 0; d < 3; ++d) out[d] = (w[d] - .5 * uw[d]) + c * b[d];
  }
}

/* ------------------------------------------------------------------ */
/* Dense inverse: LU with partial pivoting, then column solves.        */
/* ------------------------------------------------------------------ */
int or_inverse(int n, const double *A, double *Ainv) {
  double *LU = (double *)malloc(sizeof(double) * n * n);
  int *piv = (int *)malloc(sizeof(int) * n);
  memcpy(LU, A, sizeof(double) * n * n);
  int ok = 1;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(LU[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(LU[i * n + k]) > best) {
        best = fabs(LU[i * n + k]);
        p = i;
      }
    piv[k] = p;
    if (best == 0.0) {
      ok = 0;
      continue;
    }
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = LU[k * n + j];
        LU[k * n + j] = LU[p * n + j];
        LU[p * n + j] = t;
      }
    for (int i = k + 1; i < n; ++i) {
      double l = LU[i * n + k] / LU[k * n + k];
      LU[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) LU[i * n + j] -= l * LU[k * n + j];
    }
  }
  double *col = (double *)malloc(sizeof(double) * n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) col[i] = (i == c) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k) {
      int p = piv[k];
      if (p != k) {
        double t = col[k];
        col[k] = col[p];
        col[p] = t;
      }
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) col[i] -= LU[i * n + j] * col[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) col[i] -= LU[i * n + j] * col[j];
      col[i] /= LU[i * n + i];
    }
    for (int i = 0; i < n; ++i) Ainv[i * n + c] = col[i];
  }
  free(col);
  free(LU);
  free(piv);
  return ok;
}

/* ------------------------------------------------------------------ */
/* System                                                              */
/* ------------------------------------------------------------------ */
struct or_sim {
  fs_sim_params p;
  int nf, np, nlam;
  int *nw, *boff, *loff;
  long *joff;
  fs_fil_props *props;
  double *st; /* myStrainTwist per filament (3) */
  double *X, *Xs, *Xt, *U, *Us, *Ut, *q, *qs, *qt, *F, *TAU;
  double *lam, *lam1, *lam2;
  double *J, *Jinv;
  double *Cmat, *Dmat; /* maxiter columns of 6Np */
  double *Y, *Fa, *Ta, *V, *W;
  double *E, *NewE, *Upd, *tmp, *dE;
  or_colbar *cb;
  int barrier_off; /* gait-reset pass without the barrier */
  int *sfil, *sseg;
  /* spring links: global bead pairs, in link order */
  int nspring;
  int *spring; /* 2 per link */
  double spring_k, spring_len;
  /* seconds spent in the last timestep per profiler bucket (profilers.hpp:
   * 24-55): assign F, build Jinv, collision barrier, FCM, error check,
   * solve Jinv, total, then the FCM's spread (make_zero + gaussian_setup +
   * force_distribution), FFT forward, flow solve, FFT backward and
   * particle velocities/rotations */
  double tb[OR_NBUCKETS];
};

static double wall_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void or_sim_bucket_times(or_sim *h, double *out) {
  memcpy(out, h->tb, sizeof h->tb);
}

static double *dz(size_t n) { return (double *)calloc(n ? n : 1, 8); }

or_sim *or_sim_create(const fs_sim_params *p, int nfil, const int *nw,
                      const fs_fil_props *props) {
  or_sim *h = (or_sim *)calloc(1, sizeof(or_sim));
  h->p = *p;
  h->nf = nfil;
  h->nw = (int *)malloc(sizeof(int) * nfil);
  h->boff = (int *)malloc(sizeof(int) * (nfil + 1));
  h->loff = (int *)malloc(sizeof(int) * (nfil + 1));
  h->joff = (long *)malloc(sizeof(long) * (nfil + 1));
  h->props = (fs_fil_props *)malloc(sizeof(fs_fil_props) * nfil);
  memcpy(h->props, props, sizeof(fs_fil_props) * nfil);
  h->boff[0] = h->loff[0] = 0;
  h->joff[0] = 0;
  for (int f = 0; f < nfil; ++f) {
    h->nw[f] = nw[f];
    h->boff[f + 1] = h->boff[f] + nw[f];
    h->loff[f + 1] = h->loff[f] + nw[f] - 1;
    h->joff[f + 1] = h->joff[f] + (long)(6 * nw[f]) * (6 * nw[f]);
  }
  h->np = h->boff[nfil];
  h->nlam = h->loff[nfil];
  size_t np = h->np, nl = h->nlam;
  h->st = dz(3 * nfil);
  for (int f = 0; f < nfil; ++f) h->st[3 * f] = p->strain_twist0;
  h->X = dz(3 * np);
  h->Xs = dz(3 * np);
  h->Xt = dz(3 * np);
  h->U = dz(3 * np);
  h->Us = dz(3 * np);
  h->Ut = dz(3 * np);
  h->q = dz(4 * np);
  h->qs = dz(4 * np);
  h->qt = dz(4 * np);
  for (size_t i = 0; i < np; ++i) h->q[4 * i] = h->qs[4 * i] = h->qt[4 * i] = 1;
  h->F = dz(3 * np);
  h->TAU = dz(3 * np);
  h->lam = dz(3 * nl);
  h->lam1 = dz(3 * nl);
  h->lam2 = dz(3 * nl);
  h->J = dz(h->joff[nfil]);
  h->Jinv = dz(h->joff[nfil]);
  h->Cmat = dz(6 * np * (size_t)p->broyden_maxiter);
  h->Dmat = dz(6 * np * (size_t)p->broyden_maxiter);
  h->Y = dz(3 * np);
  h->Fa = dz(3 * np);
  h->Ta = dz(3 * np);
  h->V = dz(3 * np);
  h->W = dz(3 * np);
  h->E = dz(6 * np);
  h->NewE = dz(6 * np);
  h->Upd = dz(6 * np);
  h->tmp = dz(6 * np);
  h->dE = dz(6 * np);
  h->sfil = (int *)malloc(sizeof(int) * (np ? np : 1));
  h->sseg = (int *)malloc(sizeof(int) * (np ? np : 1));
  for (int f = 0; f < nfil; ++f)
    for (int i = 0; i < nw[f]; ++i) {
      h->sfil[h->boff[f] + i] = f;
      h->sseg[h->boff[f] + i] = i;
    }
  fs_colbar_params cp;
  cp.box[0] = p->fcm.nx * p->fcm.dx;
  cp.box[1] = p->fcm.ny * p->fcm.dx;
  cp.box[2] = p->fcm.nz * p->fcm.dx;
  cp.a = p->fcm.a;
  cp.fs_over_2a = p->barrier_fs_over_2a;
  cp.cap = p->barrier_cap;
  cp.initial_friction_steps = p->initial_friction_steps;
  h->cb = or_colbar_create(&cp, h->np, h->sfil, h->sseg);
  return h;
}

void or_sim_destroy(or_sim *h) {
  if (!h) return;
  double *bufs[] = {h->st,   h->X,    h->Xs,   h->Xt,  h->U,    h->Us,
                    h->Ut,   h->q,    h->qs,   h->qt,  h->F,    h->TAU,
                    h->lam,  h->lam1, h->lam2, h->J,   h->Jinv, h->Cmat,
                    h->Dmat, h->Y,    h->Fa,   h->Ta,  h->V,    h->W,
                    h->E,    h->NewE, h->Upd,  h->tmp, h->dE};
  for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; ++i) free(bufs[i]);
  free(h->nw);
  free(h->boff);
  free(h->loff);
  free(h->joff);
  free(h->props);
  free(h->sfil);
  free(h->sseg);
  or_colbar_destroy(h->cb);
  free(h);
}

void or_sim_sizes(or_sim *h, int *nfil, int *np, int *nlam) {
  *nfil = h->nf;
  *np = h->np;
  *nlam = h->nlam;
}

static void cpy(double *dst, const double *src, size_t n) {
  if (dst && src) memcpy(dst, src, sizeof(double) * n);
}

void or_sim_set_state(or_sim *h, const fs_state_ptrs *s) {
  size_t np = h->np, nl = h->nlam;
  cpy(h->X, s->X, 3 * np);
  cpy(h->Xs, s->Xs, 3 * np);
  cpy(h->Xt, s->Xt, 3 * np);
  cpy(h->U, s->U, 3 * np);
  cpy(h->Us, s->Us, 3 * np);
  cpy(h->Ut, s->Ut, 3 * np);
  cpy(h->q, s->q, 4 * np);
  cpy(h->qs, s->qs, 4 * np);
  cpy(h->qt, s->qt, 4 * np);
  cpy(h->lam, s->lam, 3 * nl);
  cpy(h->lam1, s->lam1, 3 * nl);
  cpy(h->lam2, s->lam2, 3 * nl);
  cpy(h->F, s->F, 3 * np);
  cpy(h->TAU, s->TAU, 3 * np);
}

void or_sim_get_state(or_sim *h, fs_state_ptrs *s) {
  size_t np = h->np, nl = h->nlam;
  cpy(s->X, h->X, 3 * np);
  cpy(s->Xs, h->Xs, 3 * np);
  cpy(s->Xt, h->Xt, 3 * np);
  cpy(s->U, h->U, 3 * np);
  cpy(s->Us, h->Us, 3 * np);
  cpy(s->Ut, h->Ut, 3 * np);
  cpy(s->q, h->q, 4 * np);
  cpy(s->qs, h->qs, 4 * np);
  cpy(s->qt, h->qt, 4 * np);
  cpy(s->lam, h->lam, 3 * nl);
  cpy(s->lam1, h->lam1, 3 * nl);
  cpy(s->lam2, h->lam2, 3 * nl);
  cpy(s->F, h->F, 3 * np);
  cpy(s->TAU, h->TAU, 3 * np);
}

void or_sim_fcm_arrays(or_sim *h, double **Y, double **F, double **T,
                       double **V, double **W) {
  *Y = h->Y;
  *F = h->Fa;
  *T = h->Ta;
  *V = h->V;
  *W = h->W;
}

/* unRobotArmify (Filament.hpp:489-505) */
static void unrobotarmify(or_sim *h, int f) {
  int b0 = h->boff[f], nw = h->nw[f];
  double dL = h->p.dL;
  double Xsum[3], t1[3], t2[3];
  memcpy(Xsum, h->Xs + 3 * b0, sizeof Xsum);
  or_quat_tangent(h->qs + 4 * b0, t1);
  for (int i = 1; i < nw; ++i) {
    or_quat_tangent(h->qs + 4 * (b0 + i), t2);
    for (int d = 0; d < 3; ++d) {
      Xsum[d] += (t1[d] + t2[d]) * (0.5 * dL);
      h->Xs[3 * (b0 + i) + d] = Xsum[d];
      t1[d] = t2[d];
    }
  }
}

/* Filament::initialGuess (Filament.hpp:227-237, Bead.hpp:202-217) followed
 * by unRobotArmify (main.cpp:334-336). */
void or_sim_initial_guess(or_sim *h) {
  for (int f = 0; f < h->nf; ++f) {
    int b0 = h->boff[f];
    for (int i = 0; i < h->nw[f]; ++i) {
      int g = b0 + i;
      if (i == 0)
        for (int d = 0; d < 3; ++d)
          h->Xs[3 * g + d] = 2. * h->X[3 * g + d] - h->Xt[3 * g + d];
      for (int d = 0; d < 3; ++d)
        h->Us[3 * g + d] = 2. * h->U[3 * g + d] - h->Ut[3 * g + d];
      memcpy(h->qs + 4 * g, h->q + 4 * g, 4 * sizeof(double));
      or_quat_expm(h->qs + 4 * g, h->Us + 3 * g);
    }
    unrobotarmify(h, f);
  }
}

/* applyElasticTorques (Filament.hpp:289-342) */
static void elastic_torques(or_sim *h, int f, int nt) {
  const fs_fil_props *pr = &h->props[f];
  int b0 = h->boff[f], nw = h->nw[f];
  double dL = h->p.dL;
  double tt = nt * h->p.dt;
  double Lf = (nw - 1) * dL;
  double *st = h->st + 3 * f;
  for (int i = 0; i < nw - 1; ++i) {
    int j = i + 1;
    double T0 = pr->T0;
    double s = dL * i + .5 * dL;
    double facPref;
    if (h->p.sperm_decay) {
      if (s > Lf / 2.0)
        facPref = T0 * 2.0 * (Lf - s) / Lf;
      else
        facPref = T0;
    } else {
      facPref = T0 * (Lf - s) / Lf;
    }
    const double *qi = h->qs + 4 * (b0 + i), *qj = h->qs + 4 * (b0 + j);
    double qm[4], q2[4];
    or_quat_midpoint(qi, qj, qm);
    for (int c = 0; c < 4; ++c) q2[c] = (qj[c] - qi[c]) * (1. / dL);
    q_from_left_conj(q2, qm);
    double ph = pr->wavenumber * s - pr->omega * tt + pr->phase;
    st[1] = pr->beta * facPref * cos(ph);
    st[2] = pr->alpha * facPref * sin(ph);
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = 2. * q2[1 + d] - st[d];
    v[0] *= pr->c;
    v[1] *= pr->kap;
    v[2] *= pr->kap;
    or_quat_rotate(qm, v);
    for (int d = 0; d < 3; ++d) {
      h->TAU[3 * (b0 + i) + d] += v[d];
      h->TAU[3 * (b0 + j) + d] -= v[d];
    }
  }
}

/* applyConstraintForcesTorques (Filament.hpp:345-370) */
static void constraint_forces(or_sim *h, int f) {
  int b0 = h->boff[f], l0 = h->loff[f], nw = h->nw[f];
  double fac = -.5 * h->p.dL;
  for (int i = 0; i < nw - 1; ++i) {
    int j = i + 1;
    const double *lm = h->lam + 3 * (l0 + i);
    for (int d = 0; d < 3; ++d) {
      h->F[3 * (b0 + i) + d] -= lm[d];
      h->F[3 * (b0 + j) + d] += lm[d];
    }
    double t[3], c[3];
    or_quat_tangent(h->qs + 4 * (b0 + i), t);
    cross3(t, lm, c);
    for (int d = 0; d < 3; ++d) h->TAU[3 * (b0 + i) + d] += fac * c[d];
    or_quat_tangent(h->qs + 4 * (b0 + j), t);
    cross3(t, lm, c);
    for (int d = 0; d < 3; ++d) h->TAU[3 * (b0 + j) + d] += fac * c[d];
  }
}

/* spring_links (spring_link_functions.cpp:24-45): links between bead
 * (filament A, bead B) and (filament C, bead D); the global bead pairs are
 * kept. */
int or_sim_set_spring_links(or_sim *h, int n, const int *links, double k,
                            double len) {
  free(h->spring);
  h->spring = (int *)malloc(sizeof(int) * 2 * (n > 0 ? n : 1));
  for (int l = 0; l < n; ++l)
    for (int e = 0; e < 2; ++e) {
      int f = links[4 * l + 2 * e], b = links[4 * l + 2 * e + 1];
      if (f < 0 || f >= h->nf || b < 0 || b >= h->nw[f]) return -1;
      h->spring[2 * l + e] = h->boff[f] + b;
    }
  h->nspring = n;
  h->spring_k = k;
  h->spring_len = len;
  return 0;
}

/* apply_spring_link_forces (spring_link_functions.cpp:24-45): in link
 * order, F_a += f and F_b -= f with d = VecPeriodicX(Xs_a, Xs_b)
 * (spring_link_functions.cpp:131-138), f = -k (|d| - L) d / |d|. */
static void spring_forces(or_sim *h) {
  const double Lb[3] = {h->p.fcm.nx * h->p.fcm.dx, h->p.fcm.ny * h->p.fcm.dx,
                        h->p.fcm.nz * h->p.fcm.dx};
  for (int l = 0; l < h->nspring; ++l) {
    const int a = h->spring[2 * l], b = h->spring[2 * l + 1];
    double d[3];
    for (int c = 0; c < 3; ++c) {
      d[c] = h->Xs[3 * a + c] - h->Xs[3 * b + c];
      d[c] = d[c] - Lb[c] * floor(d[c] / Lb[c] + 0.5);
    }
    const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double mag = h->spring_k * (dist - h->spring_len);
    for (int c = 0; c < 3; ++c) {
      const double f = -mag * d[c] / dist;
      h->F[3 * a + c] += f;
      h->F[3 * b + c] -= f;
    }
  }
}

/* forces + torques + assign_filament_data (main.cpp:337-361) */
void or_sim_forces(or_sim *h, int nt) {
  memset(h->F, 0, sizeof(double) * 3 * h->np);
  memset(h->TAU, 0, sizeof(double) * 3 * h->np);
  for (int f = 0; f < h->nf; ++f) {
    elastic_torques(h, f, nt);
    constraint_forces(h, f);
    /* applyExternalForcesTorques (Filament.hpp:379-402) */
    for (int i = 0; i < h->nw[f]; ++i)
      for (int d = 0; d < 3; ++d)
        h->F[3 * (h->boff[f] + i) + d] += h->p.weight_per_length[d] * h->p.dL;
  }
  spring_forces(h); /* main.cpp:354-357, 540-543 */
  /* Fcm::assign_filament_data (FCMfunctions.cpp:882-918) */
  double L[3] = {h->p.fcm.nx * h->p.fcm.dx, h->p.fcm.ny * h->p.fcm.dx,
                 h->p.fcm.nz * h->p.fcm.dx};
  for (int g = 0; g < h->np; ++g)
    for (int d = 0; d < 3; ++d) {
      h->Y[3 * g + d] = fmod(h->Xs[3 * g + d], L[d]);
      h->Fa[3 * g + d] = h->F[3 * g + d];
      h->Ta[3 * g + d] = h->TAU[3 * g + d];
    }
}

/* ---- Jacobian blocks (Filament.hpp:803-1120) ---- */
typedef double m3[3][3];

/* rcross: A v = v x x (Filament.hpp:870-882) */
static void rcross(const double *x, m3 A) {
  A[0][0] = 0.;
  A[0][1] = x[2];
  A[0][2] = -x[1];
  A[1][0] = -x[2];
  A[1][1] = 0.;
  A[1][2] = x[0];
  A[2][0] = x[1];
  A[2][1] = -x[0];
  A[2][2] = 0.;
}
static void mm(const m3 A, const m3 B, m3 C) {
  m3 T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      T[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
  memcpy(C, T, sizeof(m3));
}
static void mT(const m3 A, m3 B) {
  m3 T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[i][j] = A[j][i];
  memcpy(B, T, sizeof(m3));
}
static void mv(const m3 A, const double *v, double *o) {
  double t[3];
  for (int i = 0; i < 3; ++i)
    t[i] = A[i][0] * v[0] + A[i][1] * v[1] + A[i][2] * v[2];
  o[0] = t[0];
  o[1] = t[1];
  o[2] = t[2];
}
/* C = a*A + b*B */
static void lin2(double a, const m3 A, double b, const m3 B, m3 C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = a * A[i][j] + b * B[i][j];
}
static void outer(const double *u, const double *v, m3 C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = u[i] * v[j];
}

typedef struct {
  double t[3], n[3], b[3];
} frame;
static void get_frame(const double *q, frame *fr) {
  or_quat_tangent(q, fr->t);
  quat_normal(q, fr->n);
  quat_binormal(q, fr->b);
}

/* -2dt(dO - .5(OM + LM dO) + (rcross(OM Us) + LM OM + LM LM dO)/12)/3 */
static void diag_block(double dt, const m3 dO, const m3 LM, const m3 OM,
                       const double *Us, m3 out) {
  m3 LdO, LLdO, LOM, R;
  mm(LM, dO, LdO);
  mm(LM, LdO, LLdO);
  mm(LM, OM, LOM);
  double w[3];
  mv(OM, Us, w);
  rcross(w, R);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out[i][j] = -2 * dt *
                  (dO[i][j] - 0.5 * (OM[i][j] + LdO[i][j]) +
                   (R[i][j] + LOM[i][j] + LLdO[i][j]) / 12.) /
                  3.;
}
/* -2dt(dO - .5 LM dO + LM LM dO/12)/3 */
static void off_block(double dt, const m3 dO, const m3 LM, m3 out) {
  m3 LdO, LLdO;
  mm(LM, dO, LdO);
  mm(LM, LdO, LLdO);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out[i][j] = -2 * dt * (dO[i][j] - 0.5 * LdO[i][j] + LLdO[i][j] / 12.) / 3.;
}

/* ElasticPart pieces. sgn*0.5*C*(beta*rcross(a) + cs*0.5*(t+o) Cross/DL)
 * + sgn*0.5*KAP*(X/DL - 0.5*(ST0 rcross(n) + ST1 rcross(b))) */

/* Write a 3x3 block at (r, c) of an n x n row-major matrix. */
static void put(double *J, int n, int r, int c, const m3 B) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[(long)(r + i) * n + c + j] = B[i][j];
}
static void addb(double *J, int n, int r, int c, const m3 B) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[(long)(r + i) * n + c + j] += B[i][j];
}

static void build_jacobian_fil(or_sim *h, int f, int nt) {
  const fs_fil_props *pr = &h->props[f];
  int nw = h->nw[f], b0 = h->boff[f], l0 = h->loff[f];
  int n = 6 * nw;
  double *J = h->J + h->joff[f];
  memset(J, 0, sizeof(double) * n * n);
  double dt = h->p.dt, DL = h->p.dL, MU = h->p.fcm.mu;
  double rad = pr->radius;
  const double *ST = h->st + 3 * f;
  double myC = pr->c, myKAP = pr->kap;

  /* first-bead position rows (Filament.hpp:811-827) */
  int jj = 3 * nw + 3;
  double temp = dt / (9 * OR_PI * MU * rad);
  for (int d = 0; d < 3; ++d) {
    J[(long)d * n + d] = 1;
    J[(long)d * n + jj + d] = temp / dt;
  }
  double Jprefac = 10.;

  /* BlockC(0.5 DL) into rows 3.., cols 3.. (Filament.hpp:831, 885-909) */
  {
    double fac = 0.5 * DL;
    int R = 3 * nw - 3, Cn = 3 * nw;
    double *Cc = (double *)calloc((size_t)R * Cn, sizeof(double));
    double tk[3], tkm1[3];
    m3 A, B;
    or_quat_tangent(h->qs + 4 * (b0 + 1), tk);
    or_quat_tangent(h->qs + 4 * (b0 + 0), tkm1);
    rcross(tkm1, A);
    rcross(tk, B);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        Cc[i * Cn + j] = fac * A[i][j];
        Cc[i * Cn + 3 + j] = fac * B[i][j];
      }
    for (int k = 2; k < nw; k++) {
      int i0 = 3 * (k - 1);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < i0 + 3; ++j)
          Cc[(i0 + i) * Cn + j] = Cc[(i0 - 3 + i) * Cn + j];
      or_quat_tangent(h->qs + 4 * (b0 + k), tk);
      or_quat_tangent(h->qs + 4 * (b0 + k - 1), tkm1);
      rcross(tkm1, A);
      rcross(tk, B);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          Cc[(i0 + i) * Cn + i0 + j] += fac * A[i][j];
          Cc[(i0 + i) * Cn + i0 + 3 + j] += fac * B[i][j];
        }
    }
    for (int i = 0; i < R; ++i)
      for (int j = 0; j < Cn; ++j) J[(long)(3 + i) * n + 3 + j] = Cc[i * Cn + j];
    free(Cc);
  }

  /* Jprefac*BlockD(nt)/dt into rows 3.., cols 3nw+3.. (Filament.hpp:836,
   * 913-946) */
  for (int m = 0; m < nw - 1; ++m) {
    double facn = -2. * dt / (18. * OR_PI * MU * rad);
    if (nt <= h->p.impl_euler_steps) facn *= (3. / 2.);
    int i = 3 * m;
    for (int d = 0; d < 3; ++d)
      J[(long)(3 + i + d) * n + 3 * nw + 3 + i + d] = Jprefac * facn / dt;
    if (m < nw - 2) {
      facn = 2. * dt / (18. * OR_PI * MU * rad);
      if (nt <= h->p.impl_euler_steps) facn *= (3. / 2.);
      for (int d = 0; d < 3; ++d)
        J[(long)(3 + i + d) * n + 3 * nw + 3 + i + 3 + d] = Jprefac * facn / dt;
    }
  }

  /* BlockE into rows 3nw.., cols 3.. (Filament.hpp:841, 949-1095) */
  {
    int R0 = 3 * nw, C0 = 3;
    m3 rt, rtp1, rtm1, rn, rb, rnp1, rbp1, rnm1, rbm1, LM, OM, CP, EP, dO,
        Blk, tmp1, tmp2, rl;
    frame fm1, f0, fp1;
    double Tfac, beta, cross_part[3], lamvec[3];
    /* n = 0 */
    Tfac = 1 / (8 * OR_PI * MU * pow(rad, 3));
    get_frame(h->qs + 4 * (b0 + 0), &f0);
    get_frame(h->qs + 4 * (b0 + 1), &fp1);
    memcpy(lamvec, h->lam + 3 * (l0 + 0), sizeof lamvec);
    rcross(lamvec, rl);
    rcross(f0.t, rt);
    mm(rl, rt, CP);
    lin2(-0.5 * DL, CP, 0, CP, CP);
    beta = -ST[2] + 0.5 * (dot3(fp1.n, f0.b) - dot3(f0.n, fp1.b)) / DL;
    {
      double c1[3], c2[3];
      cross3(f0.b, fp1.n, c1);
      cross3(f0.n, fp1.b, c2);
      for (int d = 0; d < 3; ++d) cross_part[d] = c1[d] - c2[d];
    }
    double tsum[3];
    for (int d = 0; d < 3; ++d) tsum[d] = 0.5 * (f0.t[d] + fp1.t[d]);
    outer(tsum, cross_part, tmp1);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        EP[i][j] = 0.5 * myC * (tmp1[i][j] / DL + beta * rt[i][j]);
    rcross(fp1.t, rtp1);
    rcross(f0.n, rn);
    rcross(f0.b, rb);
    mm(rtp1, rt, tmp2);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        EP[i][j] += 0.5 * myKAP *
                    (tmp2[i][j] / DL - 0.5 * (ST[0] * rn[i][j] + ST[1] * rb[i][j]));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dO[i][j] = Tfac * (EP[i][j] + CP[i][j]);
    {
      m3 ru;
      rcross(h->Us + 3 * (b0 + 0), ru);
      mT(ru, LM);
      double tau[3];
      for (int d = 0; d < 3; ++d) tau[d] = Tfac * h->TAU[3 * (b0 + 0) + d];
      rcross(tau, OM);
    }
    diag_block(dt, dO, LM, OM, h->Us + 3 * (b0 + 0), Blk);
    put(J, n, R0 + 0, C0 + 0, Blk);
    /* off-diagonal (0,1) */
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        EP[i][j] = 0.5 * myC * (beta * rtp1[i][j] - tmp1[i][j] / DL);
    {
      m3 rtT;
      mT(rt, rtT);
      mm(rtT, rtp1, tmp2);
    }
    rcross(fp1.n, rnp1);
    rcross(fp1.b, rbp1);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        EP[i][j] += 0.5 * myKAP *
                    (tmp2[i][j] / DL -
                     0.5 * (ST[0] * rnp1[i][j] + ST[1] * rbp1[i][j]));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dO[i][j] = Tfac * EP[i][j];
    off_block(dt, dO, LM, Blk);
    put(J, n, R0 + 0, C0 + 3, Blk);

    /* end bead nw-1 */
    int e = nw - 1;
    Tfac = 1 / (8 * OR_PI * MU * pow(rad, 3));
    get_frame(h->qs + 4 * (b0 + e), &f0);
    get_frame(h->qs + 4 * (b0 + e - 1), &fm1);
    memcpy(lamvec, h->lam + 3 * (l0 + nw - 2), sizeof lamvec);
    rcross(lamvec, rl);
    rcross(f0.t, rt);
    mm(rl, rt, CP);
    lin2(-0.5 * DL, CP, 0, CP, CP);
    beta = -ST[2] + 0.5 * (dot3(f0.n, fm1.b) - dot3(fm1.n, f0.b)) / DL;
    {
      double c1[3], c2[3];
      cross3(fm1.b, f0.n, c1);
      cross3(fm1.n, f0.b, c2);
      for (int d = 0; d < 3; ++d) cross_part[d] = c1[d] - c2[d];
    }
    for (int d = 0; d < 3; ++d) tsum[d] = 0.5 * (f0.t[d] + fm1.t[d]);
    outer(tsum, cross_part, tmp1);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        EP[i][j] = -0.5 * myC * (beta * rt[i][j] - tmp1[i][j] / DL);
    rcross(fm1.t, rtm1);
    rcross(f0.n, rn);
    rcross(f0.b, rb);
    {
      m3 rtm1T;
      mT(rtm1, rtm1T);
      mm(rtm1T, rt, tmp2);
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        EP[i][j] -= 0.5 * myKAP *
                    (tmp2[i][j] / DL - 0.5 * (ST[0] * rn[i][j] + ST[1] * rb[i][j]));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dO[i][j] = Tfac * (EP[i][j] + CP[i][j]);
    {
      m3 ru;
      rcross(h->Us + 3 * (b0 + e), ru);
      mT(ru, LM);
      double tau[3];
      for (int d = 0; d < 3; ++d) tau[d] = Tfac * h->TAU[3 * (b0 + e) + d];
      rcross(tau, OM);
    }
    diag_block(dt, dO, LM, OM, h->Us + 3 * (b0 + e), Blk);
    put(J, n, R0 + 3 * e, C0 + 3 * e, Blk);
    /* off-diagonal (e, e-1) */
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        EP[i][j] = -0.5 * myC * (beta * rtm1[i][j] - tmp1[i][j] / DL);
    mm(rt, rtm1, tmp2);
    rcross(fm1.n, rnm1);
    rcross(fm1.b, rbm1);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        EP[i][j] -= 0.5 * myKAP *
                    (tmp2[i][j] / DL -
                     0.5 * (ST[0] * rnm1[i][j] + ST[1] * rbm1[i][j]));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dO[i][j] = Tfac * EP[i][j];
    off_block(dt, dO, LM, Blk);
    put(J, n, R0 + 3 * e, C0 + 3 * e - 3, Blk);

    /* interior beads */
    for (int k = 1; k < nw - 1; k++) {
      Tfac = 1 / (8 * OR_PI * MU * pow(rad, 3.));
      get_frame(h->qs + 4 * (b0 + k - 1), &fm1);
      get_frame(h->qs + 4 * (b0 + k), &f0);
      get_frame(h->qs + 4 * (b0 + k + 1), &fp1);
      {
        m3 ru;
        rcross(h->Us + 3 * (b0 + k), ru);
        mT(ru, LM);
        double tau[3];
        for (int d = 0; d < 3; ++d) tau[d] = Tfac * h->TAU[3 * (b0 + k) + d];
        rcross(tau, OM);
      }
      for (int d = 0; d < 3; ++d)
        lamvec[d] = h->lam[3 * (l0 + k - 1) + d] + h->lam[3 * (l0 + k) + d];
      rcross(lamvec, rl);
      rcross(f0.t, rt);
      mm(rl, rt, CP);
      lin2(-0.5 * DL, CP, 0, CP, CP);
      beta = -ST[2] + 0.5 * (dot3(fp1.n, f0.b) - dot3(f0.n, fp1.b)) / DL;
      {
        double c1[3], c2[3];
        cross3(f0.b, fp1.n, c1);
        cross3(f0.n, fp1.b, c2);
        for (int d = 0; d < 3; ++d) cross_part[d] = c1[d] - c2[d];
      }
      /* forward neighbour terms */
      for (int d = 0; d < 3; ++d) tsum[d] = 0.5 * (f0.t[d] + fp1.t[d]);
      outer(tsum, cross_part, tmp1);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          EP[i][j] = 0.5 * myC * (tmp1[i][j] / DL + beta * rt[i][j]);
      rcross(fp1.t, rtp1);
      rcross(f0.n, rn);
      rcross(f0.b, rb);
      mm(rtp1, rt, tmp2);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          EP[i][j] += 0.5 * myKAP *
                      (tmp2[i][j] / DL -
                       0.5 * (ST[0] * rn[i][j] + ST[1] * rb[i][j]));
      /* backward neighbour terms: beta is recomputed, the cross part is
       * reused from the forward side (as in Filament.hpp:1065-1067) */
      beta = -ST[2] + 0.5 * (dot3(f0.n, fm1.b) - dot3(fm1.n, f0.b)) / DL;
      double tsm[3];
      for (int d = 0; d < 3; ++d) tsm[d] = 0.5 * (f0.t[d] + fm1.t[d]);
      m3 tmp3;
      outer(tsm, cross_part, tmp3);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          EP[i][j] -= 0.5 * myC * (beta * rt[i][j] - tmp3[i][j] / DL);
      rcross(fm1.t, rtm1);
      {
        m3 rtm1T;
        mT(rtm1, rtm1T);
        mm(rtm1T, rt, tmp2);
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          EP[i][j] -= 0.5 * myKAP *
                      (tmp2[i][j] / DL -
                       0.5 * (ST[0] * rn[i][j] + ST[1] * rb[i][j]));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) dO[i][j] = Tfac * (CP[i][j] + EP[i][j]);
      diag_block(dt, dO, LM, OM, h->Us + 3 * (b0 + k), Blk);
      put(J, n, R0 + 3 * k, C0 + 3 * k, Blk);
      /* left off-diagonal */
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          EP[i][j] = -0.5 * myC * (beta * rtm1[i][j] - tmp3[i][j] / DL);
      mm(rt, rtm1, tmp2);
      rcross(fm1.n, rnm1);
      rcross(fm1.b, rbm1);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          EP[i][j] -= 0.5 * myKAP *
                      (tmp2[i][j] / DL -
                       0.5 * (ST[0] * rnm1[i][j] + ST[1] * rbm1[i][j]));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) dO[i][j] = Tfac * EP[i][j];
      off_block(dt, dO, LM, Blk);
      put(J, n, R0 + 3 * k, C0 + 3 * k - 3, Blk);
      /* right off-diagonal (beta is the backward one here) */
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          EP[i][j] = 0.5 * myC * (beta * rtp1[i][j] - tmp1[i][j] / DL);
      {
        m3 rtT;
        mT(rt, rtT);
        mm(rtT, rtp1, tmp2);
      }
      rcross(fp1.n, rnp1);
      rcross(fp1.b, rbp1);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          EP[i][j] += 0.5 * myKAP *
                      (tmp2[i][j] / DL -
                       0.5 * (ST[0] * rnp1[i][j] + ST[1] * rbp1[i][j]));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) dO[i][j] = Tfac * EP[i][j];
      off_block(dt, dO, LM, Blk);
      put(J, n, R0 + 3 * k, C0 + 3 * k + 3, Blk);
    }
    for (int k = 0; k < 3 * nw; k++) J[(long)(R0 + k) * n + C0 + k] += 1.;
  }

  /* BlockF/dt into rows 3nw.., cols 3nw+3.. (Filament.hpp:853, 1098-1120) */
  for (int i = 0; i < nw - 1; i++) {
    int j = 3 * i;
    double fac = dt * DL / (24 * OR_PI * MU * pow(rad, 3.));
    m3 tm, tmat, umat, um_tm, uu_tm, Blk;
    double tv[3];
    or_quat_tangent(h->qs + 4 * (b0 + i), tv);
    rcross(tv, tm);
    mT(tm, tmat);
    rcross(h->Us + 3 * (b0 + i), umat);
    mm(umat, tmat, um_tm);
    mm(umat, um_tm, uu_tm);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        Blk[a][b] =
            fac * (tmat[a][b] - 0.5 * um_tm[a][b] + uu_tm[a][b] / 12.) / dt;
    addb(J, n, 3 * nw + j, 3 * nw + 3 + j, Blk);
    fac = dt * DL / (24. * OR_PI * MU * pow(rad, 3));
    or_quat_tangent(h->qs + 4 * (b0 + i + 1), tv);
    rcross(tv, tm);
    mT(tm, tmat);
    rcross(h->Us + 3 * (b0 + i + 1), umat);
    mm(umat, tmat, um_tm);
    mm(umat, um_tm, uu_tm);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        Blk[a][b] =
            fac * (tmat[a][b] - 0.5 * um_tm[a][b] + uu_tm[a][b] / 12.) / dt;
    addb(J, n, 3 * nw + j + 3, 3 * nw + 3 + j, Blk);
  }

  /* inv(J) and the lambda-row rescale (Filament.hpp:862-865) */
  double *Ji = h->Jinv + h->joff[f];
  or_inverse(n, J, Ji);
  for (int r = 3 * nw + 3; r < 6 * nw; ++r)
    for (int c = 0; c < n; ++c) Ji[(long)r * n + c] /= dt;
}

void or_sim_build_jacobian(or_sim *h, int nt) {
  memset(h->Cmat, 0, sizeof(double) * 6 * (size_t)h->np * h->p.broyden_maxiter);
  memset(h->Dmat, 0, sizeof(double) * 6 * (size_t)h->np * h->p.broyden_maxiter);
  /* filaments are independent: one thread per filament at a time */
#pragma omp parallel for schedule(dynamic, 4)
  for (int f = 0; f < h->nf; ++f) build_jacobian_fil(h, f, nt);
}

void or_sim_get_jacobian(or_sim *h, int f, double *out) {
  long n = 6 * h->nw[f];
  memcpy(out, h->J + h->joff[f], sizeof(double) * n * n);
}

void or_sim_get_jinv(or_sim *h, int f, double *out) {
  long n = 6 * h->nw[f];
  memcpy(out, h->Jinv + h->joff[f], sizeof(double) * n * n);
}

/* filamentJsolver::solve (FilamentJacobianSolver.hpp:49-73) */
void or_sim_jsolve(or_sim *h, const double *b, double *x, int iter) {
  size_t N6 = 6 * (size_t)h->np;
  double *xo = (double *)malloc(sizeof(double) * N6);
#pragma omp parallel for schedule(dynamic, 8)
  for (int f = 0; f < h->nf; ++f) {
    int n = 6 * h->nw[f];
    size_t so = 6 * (size_t)h->boff[f];
    const double *Ji = h->Jinv + h->joff[f];
    for (int r = 0; r < n; ++r) {
      double s = 0;
      for (int c = 0; c < n; ++c) s += Ji[(long)r * n + c] * b[so + c];
      xo[so + r] = s;
    }
  }
  for (int i = 0; i < iter; ++i) {
    const double *D = h->Dmat + N6 * i, *C = h->Cmat + N6 * i;
    double d = 0;
    for (size_t r = 0; r < N6; ++r) d += D[r] * b[r];
    for (size_t r = 0; r < N6; ++r) xo[r] += C[r] * d;
  }
  memcpy(x, xo, sizeof(double) * N6);
  free(xo);
}

/* bad-Broyden column (main.cpp:640-663) */
void or_sim_broyden_update(or_sim *h, const double *E, const double *Enew,
                           const double *update, int iter) {
  size_t N6 = 6 * (size_t)h->np;
  double nrm = 0;
  for (size_t r = 0; r < N6; ++r) {
    h->dE[r] = Enew[r] - E[r];
    nrm += h->dE[r] * h->dE[r];
  }
  nrm = sqrt(nrm);
  or_sim_jsolve(h, h->dE, h->tmp, iter);
  double *C = h->Cmat + N6 * iter, *D = h->Dmat + N6 * iter;
  for (size_t r = 0; r < N6; ++r) {
    double t = h->tmp[r] * (-1. / nrm);
    t -= update[r] / nrm;
    C[r] = t;
    D[r] = h->dE[r] / nrm;
  }
}

/* Filament::update (Filament.hpp:425-452, Bead.hpp:236-256) +
 * unRobotArmify (main.cpp:528-531) */
void or_sim_update(or_sim *h, const double *u) {
  for (int f = 0; f < h->nf; ++f) {
    int b0 = h->boff[f], l0 = h->loff[f], nw = h->nw[f];
    size_t j = 6 * (size_t)b0;
    for (int d = 0; d < 3; ++d) {
      h->Xs[3 * b0 + d] -= u[j + d];
      h->Us[3 * b0 + d] -= u[j + 3 + d];
    }
    memcpy(h->qs + 4 * b0, h->q + 4 * b0, 4 * sizeof(double));
    or_quat_expm(h->qs + 4 * b0, h->Us + 3 * b0);
    j += 6;
    for (int i = 1; i < nw; ++i) {
      int g = b0 + i;
      for (int d = 0; d < 3; ++d) h->Us[3 * g + d] -= u[j + d];
      memcpy(h->qs + 4 * g, h->q + 4 * g, 4 * sizeof(double));
      or_quat_expm(h->qs + 4 * g, h->Us + 3 * g);
      j += 3;
    }
    for (int i = 0; i < nw - 1; ++i) {
      for (int d = 0; d < 3; ++d) h->lam[3 * (l0 + i) + d] -= u[j + d];
      j += 3;
    }
    unrobotarmify(h, f);
  }
}

/* ErrorCheckFCM (MobilityConstraints.cpp:447-770). Returns check. */
int or_sim_residual(or_sim *h, int nt, const double *V, const double *W,
                    double *err, double *maxabs) {
  int check = 0;
  double mx = 0.0;
  double dt = h->p.dt, TOL = h->p.tol;
  int euler = nt <= h->p.impl_euler_steps;
  const double fourthirds = 4. / 3., onethird = 1. / 3., twothirdsDt = dt * 2. / 3.;
  for (int f = 0; f < h->nf; ++f) {
    int b0 = h->boff[f], nw = h->nw[f];
    size_t k0 = 6 * (size_t)b0;
    for (int i = 0; i < nw; ++i) {
      int g = b0 + i;
      double e1[3], e2[3], dw[3];
      or_dexpinv(h->Us + 3 * g, W + 3 * g, dw);
      for (int d = 0; d < 3; ++d) {
        if (euler) {
          e1[d] = h->Xs[3 * g + d] - h->X[3 * g + d] - dt * V[3 * g + d];
          e2[d] = h->Us[3 * g + d] - dt * dw[d];
        } else {
          e1[d] = h->Xs[3 * g + d] - fourthirds * h->X[3 * g + d] +
                  onethird * h->Xt[3 * g + d] - twothirdsDt * V[3 * g + d];
          e2[d] = h->Us[3 * g + d] - onethird * h->U[3 * g + d] -
                  twothirdsDt * dw[d];
        }
        err[k0 + 3 * i + d] = e1[d];
        err[k0 + 3 * nw + 3 * i + d] = e2[d];
        if (fabs(e1[d]) > TOL || fabs(e2[d]) > TOL) check = 1;
        if (fabs(e1[d]) > mx) mx = fabs(e1[d]);
        if (fabs(e2[d]) > mx) mx = fabs(e2[d]);
      }
    }
  }
  if (maxabs) *maxabs = mx;
  return check;
}

/* Filament::step (Filament.hpp:455-478, Bead.hpp:266-276) */
void or_sim_step(or_sim *h, int nt) {
  for (int g = 0; g < h->np; ++g) {
    for (int d = 0; d < 3; ++d) {
      h->Xt[3 * g + d] = h->X[3 * g + d];
      h->X[3 * g + d] = h->Xs[3 * g + d];
      h->Ut[3 * g + d] = h->U[3 * g + d];
      h->U[3 * g + d] = h->Us[3 * g + d];
    }
    memcpy(h->qt + 4 * g, h->q + 4 * g, 4 * sizeof(double));
    or_quat_expm(h->q + 4 * g, h->U + 3 * g);
  }
  for (int l = 0; l < 3 * h->nlam; ++l) {
    if (nt > 12) {
      double t = 3 * h->lam[l] - 3 * h->lam1[l] + h->lam2[l];
      h->lam2[l] = h->lam1[l];
      h->lam1[l] = h->lam[l];
      h->lam[l] = t;
    } else if (nt == 12) {
      double t = 2. * h->lam[l] - h->lam1[l];
      h->lam2[l] = h->lam1[l];
      h->lam1[l] = h->lam[l];
      h->lam[l] = t;
    } else {
      h->lam1[l] = h->lam[l];
    }
  }
}

void or_sim_mobility(or_sim *h, int nt) {
  if (nt > h->p.initial_friction_steps)
    or_fcm_mobility(&h->p.fcm, h->np, h->Y, h->Fa, h->Ta, h->V, h->W);
  else
    or_fcm_friction(&h->p.fcm, h->np, h->Fa, h->Ta, h->V, h->W);
}

double or_sim_barrier(or_sim *h, int nt, int link) {
  if (link) or_colbar_link(h->cb, h->Xs);
  return or_colbar_apply(h->cb, h->Y, nt, h->Fa);
}

void or_sim_barrier_force(or_sim *h, double *out) {
  if (h->cb)
    or_colbar_last_force(h->cb, out);
  else
    memset(out, 0, sizeof(double) * 3 * (size_t)h->np);
}

void or_sim_set_barrier(or_sim *h, int on, int mask) {
  h->barrier_off = !on;
  or_colbar_set_mask(h->cb, mask);
}

/* main.cpp:284-754; commit = 0: a trial pass of main_gaitreset.cpp
 * (gait_reset_i < 2), which skips Filament::step (:775-781) */
/* the FCM solve, its time and its stage times into the buckets */
static void timed_mobility(or_sim *h, int nt) {
  double t0 = wall_s();
  or_sim_mobility(h, nt);
  h->tb[3] += wall_s() - t0;
  if (nt > h->p.initial_friction_steps) {
    double st[5];
    or_fcm_last_times(st);
    for (int k = 0; k < 5; ++k) h->tb[7 + k] += st[k];
  }
}

static int timestep_impl(or_sim *h, int nt, int commit) {
  size_t N6 = 6 * (size_t)h->np;
  const int bar = h->p.barrier_on && !h->barrier_off;
  memset(h->tb, 0, sizeof h->tb);
  double t0 = wall_s(), t;
  or_sim_initial_guess(h);
  or_sim_forces(h, nt);
  h->tb[0] += (t = wall_s()) - t0;
  or_sim_build_jacobian(h, nt);
  h->tb[1] += wall_s() - t;
  t = wall_s();
  if (bar) or_sim_barrier(h, nt, 1);
  h->tb[2] += wall_s() - t;
  timed_mobility(h, nt);
  t = wall_s();
  int check = or_sim_residual(h, nt, h->V, h->W, h->E, NULL);
  h->tb[4] += wall_s() - t;
  int iter = 0;
  while (check && iter < h->p.broyden_maxiter) {
    t = wall_s();
    or_sim_jsolve(h, h->E, h->Upd, iter);
    h->tb[5] += wall_s() - t;
    t = wall_s();
    or_sim_update(h, h->Upd);
    or_sim_forces(h, nt);
    h->tb[0] += wall_s() - t;
    t = wall_s();
    if (bar) or_sim_barrier(h, nt, 0);
    h->tb[2] += wall_s() - t;
    timed_mobility(h, nt);
    t = wall_s();
    check = or_sim_residual(h, nt, h->V, h->W, h->NewE, NULL);
    h->tb[4] += wall_s() - t;
    t = wall_s();
    or_sim_broyden_update(h, h->E, h->NewE, h->Upd, iter);
    memcpy(h->E, h->NewE, sizeof(double) * N6);
    h->tb[5] += wall_s() - t;
    ++iter;
  }
  if (commit) or_sim_step(h, nt);
  h->tb[6] = wall_s() - t0;
  return iter;
}

int or_sim_timestep(or_sim *h, int nt) { return timestep_impl(h, nt, 1); }
int or_sim_timestep_trial(or_sim *h, int nt) {
  return timestep_impl(h, nt, 0);
}
