// filament.hip -- filament mechanics, BDF2 residual, approximate Jacobian,
// bad-Broyden solve and the implicit timestep for gfx950 (FP64 throughout).
//
// Replaces class Filament (Filament.hpp:21-1123), class Bead (Bead.hpp),
// UnitQuaternion (UnitQuaternion.hpp), ErrorCheckFCM
// (MobilityConstraints.cpp:447-770), filamentJsolver
// (FilamentJacobianSolver.hpp) and the Broyden loop of main.cpp:284-754.
// State is structure-of-arrays in HBM, indexed by global bead g (per bead)
// and global link l (per link; links of filament f start at loff[f]).
// Kernels:
//   k_guess       initialGuess + unRobotArmify (one workgroup per filament)
//   k_update      update(u) + unRobotArmify
//   k_forces      elastic torques, constraint forces/torques, gravity and
//                 assign_filament_data (Y = fmod(Xs, L), F, T)
//   k_residual    ErrorCheckFCM per bead + exact infinity-norm reduction
//   k_jac_build   per-filament approximate Jacobian (rcross, BlockC/D/E/F)
//   k_gj_inverse  Gauss-Jordan with partial pivoting, one workgroup per
//                 filament; writes J^-1 column-major with the lambda-row
//                 rescale of Filament.hpp:863-865
//   k_jsolve      block-diagonal J^-1 b, then the rank-one history
//                 x += sum_i C_i (D_i . b) with deterministic reductions
//   k_step        Filament::step + Bead::step
// Every kernel here is compiled with FP contraction off, so each product
// rounds as in the reference's x86-64 build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "colbar_impl.hpp"
#include "common.hpp"
#include "fcm_impl.hpp"
#include "system_impl.hpp"

#pragma clang fp contract(off)

namespace fs {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kMaxNw = 256;

// ---------------------------------------------------------------------
// quaternion algebra (UnitQuaternion.hpp), q = [scal, v0, v1, v2]
// ---------------------------------------------------------------------
__device__ __forceinline__ void dcross(const double *a, const double *b,
                                       double *c) {
  double c0 = a[1] * b[2] - a[2] * b[1];
  double c1 = a[2] * b[0] - a[0] * b[2];
  double c2 = a[0] * b[1] - a[1] * b[0];
  c[0] = c0;
  c[1] = c1;
  c[2] = c2;
}
__device__ __forceinline__ double ddot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// p * q in place on q (quaproFromLeftInPlace, UnitQuaternion.hpp:137-143)
__device__ __forceinline__ void q_from_left(double *q, const double *p) {
  double s = q[0] * p[0] - ddot(q + 1, p + 1);
  double t[3];
  dcross(p + 1, q + 1, t);
  double v0 = q[0] * p[1] + p[0] * q[1] + t[0];
  double v1 = q[0] * p[2] + p[0] * q[2] + t[1];
  double v2 = q[0] * p[3] + p[0] * q[3] + t[2];
  q[0] = s;
  q[1] = v0;
  q[2] = v1;
  q[3] = v2;
}
// (p^*) * q in place (quaproFromLeftInPlaceConj, :146-153)
__device__ __forceinline__ void q_from_left_conj(double *q, const double *p) {
  double s = q[0] * p[0] + ddot(q + 1, p + 1);
  double t[3];
  dcross(q + 1, p + 1, t);
  double v0 = -q[0] * p[1] + p[0] * q[1] + t[0];
  double v1 = -q[0] * p[2] + p[0] * q[2] + t[1];
  double v2 = -q[0] * p[3] + p[0] * q[3] + t[2];
  q[0] = s;
  q[1] = v0;
  q[2] = v1;
  q[3] = v2;
}
// q * p in place (quaproInPlace, :119-125)
__device__ __forceinline__ void q_times(double *q, const double *p) {
  double s = q[0] * p[0] - ddot(q + 1, p + 1);
  double t[3];
  dcross(q + 1, p + 1, t);
  double v0 = q[0] * p[1] + p[0] * q[1] + t[0];
  double v1 = q[0] * p[2] + p[0] * q[2] + t[1];
  double v2 = q[0] * p[3] + p[0] * q[3] + t[2];
  q[0] = s;
  q[1] = v0;
  q[2] = v1;
  q[3] = v2;
}
// expmInPlace (:227-242): q <- exp(v) q
__device__ __forceinline__ void q_expm(double *q, const double *v) {
  double theta = sqrt(ddot(v, v));
  if (theta > 1.e-12) {
    double tmp = sin(theta * .5) / theta;
    double qu[4] = {cos(theta * .5), tmp * v[0], tmp * v[1], tmp * v[2]};
    q_from_left(q, qu);
  }
}
// assignMidPointOf (:472-489) with quaSquareRootTimInPlace (:411-421)
__device__ __forceinline__ void q_midpoint(const double *p1, const double *p2,
                                           double *out) {
  double t[3];
  dcross(p2 + 1, p1 + 1, t);
  double s = p2[0] * p1[0] + ddot(p2 + 1, p1 + 1);
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
// tangent / normal / binormal (:341-378)
__device__ __forceinline__ void q_tangent(const double *q, double *v) {
  double q22 = 2 * q[2] * q[2];
  double q33 = 2 * q[3] * q[3];
  v[0] = (1 - q22 - q33);
  v[1] = 2 * (q[1] * q[2] + q[0] * q[3]);
  v[2] = 2 * (q[1] * q[3] - q[0] * q[2]);
}
__device__ __forceinline__ void q_normal(const double *q, double *v) {
  double q11 = 2 * q[1] * q[1];
  double q33 = 2 * q[3] * q[3];
  v[0] = 2 * (q[1] * q[2] - q[0] * q[3]);
  v[1] = (1 - q11 - q33);
  v[2] = 2 * (q[2] * q[3] + q[0] * q[1]);
}
__device__ __forceinline__ void q_binormal(const double *q, double *v) {
  double q11 = 2 * q[1] * q[1];
  double q22 = 2 * q[2] * q[2];
  v[0] = 2 * (q[1] * q[3] + q[0] * q[2]);
  v[1] = 2 * (q[2] * q[3] - q[0] * q[1]);
  v[2] = (1 - q11 - q22);
}
// RotateVector (:323-338)
__device__ __forceinline__ void q_rotate(const double *q, double *v) {
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
// dexpinv (MobilityConstraints.cpp:109-143)
__device__ __forceinline__ void d_dexpinv(const double *u, const double *w,
                                          double *out) {
  double tmp = sqrt(ddot(u, u));
  double uw[3];
  dcross(u, w, uw);
  if (tmp >= 1.e-12) {
    double tmp2 = tmp / sin(.5 * tmp);
    tmp2 *= cos(.5 * tmp);
    double c = .5 * tmp2 - 1.;
    double un[3] = {u[0] / tmp, u[1] / tmp, u[2] / tmp};
    double a[3], b[3];
    dcross(un, w, a);
    dcross(un, a, b);
    for (int d = 0; d < 3; ++d) out[d] = (w[d] - .5 * uw[d]) - c * b[d];
  } else {
    double c = 1. / 12;
    double b[3];
    dcross(u, uw, b);
    for (int d = 0; d < 3; ++d) out[d] = (w[d] - .5 * uw[d]) + c * b[d];
  }
}

// ---------------------------------------------------------------------
// kinematics
// ---------------------------------------------------------------------
// unRobotArmify (Filament.hpp:489-505), run by one lane.
// Called by the whole workgroup after qs is final: the tangents are formed
// in parallel into LDS, then one thread accumulates the positions in bead
// order (the same operations and order as a serial loop).
__device__ void unrobotarmify(const SysDev &S, int b0, int nw) {
  __shared__ double s_tan[kMaxNw][3];
  for (int i = threadIdx.x; i < nw; i += blockDim.x)
    q_tangent(S.qs + 4 * (b0 + i), s_tan[i]);
  __syncthreads();
  if (threadIdx.x == 0) {  // no early return: callers go on after it
    double Xsum[3] = {S.Xs[3 * b0], S.Xs[3 * b0 + 1], S.Xs[3 * b0 + 2]};
    for (int i = 1; i < nw; ++i)
      for (int d = 0; d < 3; ++d) {
        Xsum[d] += (s_tan[i - 1][d] + s_tan[i][d]) * (0.5 * S.dL);
        S.Xs[3 * (b0 + i) + d] = Xsum[d];
      }
  }
}

// initialGuess (Filament.hpp:227-237, Bead.hpp:202-217) + unRobotArmify
__global__ void k_guess(SysDev S) {
  const int f = blockIdx.x;
  const int b0 = S.boff[f], nw = S.nw[f];
  for (int i = threadIdx.x; i < nw; i += blockDim.x) {
    int g = b0 + i;
    if (i == 0)
      for (int d = 0; d < 3; ++d)
        S.Xs[3 * g + d] = 2. * S.X[3 * g + d] - S.Xt[3 * g + d];
    for (int d = 0; d < 3; ++d)
      S.Us[3 * g + d] = 2. * S.U[3 * g + d] - S.Ut[3 * g + d];
    double q[4] = {S.q[4 * g], S.q[4 * g + 1], S.q[4 * g + 2], S.q[4 * g + 3]};
    q_expm(q, S.Us + 3 * g);
    for (int c = 0; c < 4; ++c) S.qs[4 * g + c] = q[c];
  }
  __syncthreads();
  unrobotarmify(S, b0, nw);
}

// Filament::update (Filament.hpp:425-452, Bead.hpp:236-256) + unRobotArmify
__device__ __forceinline__ void d_update(const SysDev &S, const double *__restrict__ u) {
  const int f = blockIdx.x;
  const int b0 = S.boff[f], l0 = S.loff[f], nw = S.nw[f];
  const size_t so = 6 * (size_t)b0;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) {
    int g = b0 + i;
    if (i == 0) {
      for (int d = 0; d < 3; ++d) {
        S.Xs[3 * g + d] -= u[so + d];
        S.Us[3 * g + d] -= u[so + 3 + d];
      }
    } else {
      for (int d = 0; d < 3; ++d) S.Us[3 * g + d] -= u[so + 6 + 3 * (i - 1) + d];
    }
    double q[4] = {S.q[4 * g], S.q[4 * g + 1], S.q[4 * g + 2], S.q[4 * g + 3]};
    q_expm(q, S.Us + 3 * g);
    for (int c = 0; c < 4; ++c) S.qs[4 * g + c] = q[c];
  }
  for (int l = threadIdx.x; l < nw - 1; l += blockDim.x)
    for (int d = 0; d < 3; ++d)
      S.lam[3 * (l0 + l) + d] -= u[so + 3 * nw + 3 + 3 * l + d];
  __syncthreads();
  unrobotarmify(S, b0, nw);
}
__global__ void k_update(SysDev S, const double *__restrict__ u) {
  if (S.act && !*S.act) return;
  d_update(S, u);
}

// setZeroForcesTorques + applyElasticTorques + applyConstraintForcesTorques
// + applyExternalForcesTorques (Filament.hpp:271-402) + assign_filament_data
// (FCMfunctions.cpp:882-918). One workgroup per filament.
__device__ __forceinline__ void d_forces(const SysDev &S, int nt) {
  if (S.ntp) nt = *S.ntp;
  __shared__ double s_v[kMaxNw][3];  // elastic torque of link l
  __shared__ double s_t[kMaxNw][3];  // tangent of bead i
  const int f = blockIdx.x;
  const int b0 = S.boff[f], l0 = S.loff[f], nw = S.nw[f];
  const fs_fil_props pr = S.props[f];
  const double dL = S.dL;
  const double tt = nt * S.dt;
  const double Lf = (nw - 1) * dL;
  for (int i = threadIdx.x; i < nw - 1; i += blockDim.x) {
    double T0 = pr.T0;
    double s = dL * i + .5 * dL;
    double facPref;
    if (S.sperm_decay) {
      if (s > Lf / 2.0)
        facPref = T0 * 2.0 * (Lf - s) / Lf;
      else
        facPref = T0;
    } else {
      facPref = T0 * (Lf - s) / Lf;
    }
    const double *qi = S.qs + 4 * (b0 + i), *qj = S.qs + 4 * (b0 + i + 1);
    double qm[4], q2[4];
    q_midpoint(qi, qj, qm);
    for (int c = 0; c < 4; ++c) q2[c] = (qj[c] - qi[c]) * (1. / dL);
    q_from_left_conj(q2, qm);
    double ph = pr.wavenumber * s - pr.omega * tt + pr.phase;
    double st1 = pr.beta * facPref * cos(ph);
    double st2 = pr.alpha * facPref * sin(ph);
    double v[3] = {2. * q2[1] - S.st0, 2. * q2[2] - st1, 2. * q2[3] - st2};
    v[0] *= pr.c;
    v[1] *= pr.kap;
    v[2] *= pr.kap;
    q_rotate(qm, v);
    s_v[i][0] = v[0];
    s_v[i][1] = v[1];
    s_v[i][2] = v[2];
    if (i == nw - 2) {  // myStrainTwist keeps the last link's values
      S.st[3 * f + 0] = S.st0;
      S.st[3 * f + 1] = st1;
      S.st[3 * f + 2] = st2;
    }
  }
  for (int i = threadIdx.x; i < nw; i += blockDim.x)
    q_tangent(S.qs + 4 * (b0 + i), s_t[i]);
  __syncthreads();
  const double fac = -.5 * dL;
  const double Lb[3] = {S.Lx, S.Ly, S.Lz};
  for (int k = threadIdx.x; k < nw; k += blockDim.x) {
    int g = b0 + k;
    double tau[3] = {0, 0, 0}, F[3] = {0, 0, 0};
    // elastic: link k-1 (as j) then link k (as i)
    if (k > 0)
      for (int d = 0; d < 3; ++d) tau[d] -= s_v[k - 1][d];
    if (k < nw - 1)
      for (int d = 0; d < 3; ++d) tau[d] += s_v[k][d];
    // constraint: link k-1 (as j) then link k (as i)
    double c[3];
    if (k > 0) {
      const double *lm = S.lam + 3 * (l0 + k - 1);
      for (int d = 0; d < 3; ++d) F[d] += lm[d];
      dcross(s_t[k], lm, c);
      for (int d = 0; d < 3; ++d) tau[d] += fac * c[d];
    }
    if (k < nw - 1) {
      const double *lm = S.lam + 3 * (l0 + k);
      for (int d = 0; d < 3; ++d) F[d] -= lm[d];
      dcross(s_t[k], lm, c);
      for (int d = 0; d < 3; ++d) tau[d] += fac * c[d];
    }
    for (int d = 0; d < 3; ++d) {
      F[d] += S.weight[d] * dL;
      S.F[3 * g + d] = F[d];
      S.TAU[3 * g + d] = tau[d];
      S.Ya[3 * g + d] = fmod(S.Xs[3 * g + d], Lb[d]);
      S.Fa[3 * g + d] = F[d];
      S.Ta[3 * g + d] = tau[d];
    }
  }
}
__global__ void k_forces(SysDev S, int nt) {
  if (S.act && !*S.act) return;
  d_forces(S, nt);
}
// Filament::update then the forces of the updated state, one workgroup per
// filament: both read and write only their own filament's beads
__global__ void k_update_forces(SysDev S, const double *__restrict__ u, int nt) {
  if (S.act && !*S.act) return;
  d_update(S, u);
  __syncthreads();
  d_forces(S, nt);
}

// Spring links (spring_links::apply_spring_link_forces,
// spring_link_functions.cpp:24-45), called after the filament forces
// (main.cpp:354-357, 540-543) and before assign_filament_data. One thread
// per linked bead adds its links' forces in link order, the order in which
// the reference accumulates them: +f on the first bead of a link, -f on the
// second, f = -k (|d| - L) d / |d| with d = VecPeriodicX(Xs_a, Xs_b)
// (spring_link_functions.cpp:131-138). ent[e] = 2 * link + (second bead).
__global__ void k_springs(SysDev S, int nb, const int *__restrict__ bead,
                          const int *__restrict__ start,
                          const int *__restrict__ ent,
                          const int *__restrict__ pair, double k, double len) {
#pragma clang fp contract(off)
  if (S.act && !*S.act) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int g = bead[i];
  const double Lb[3] = {S.Lx, S.Ly, S.Lz};
  double F[3] = {S.F[3 * g], S.F[3 * g + 1], S.F[3 * g + 2]};
  for (int e = start[i]; e < start[i + 1]; ++e) {
    const int l = ent[e] >> 1, second = ent[e] & 1;
    const int a = pair[2 * l], b = pair[2 * l + 1];
    double dv[3];
    for (int c = 0; c < 3; ++c) {
      dv[c] = S.Xs[3 * a + c] - S.Xs[3 * b + c];
      dv[c] = dv[c] - Lb[c] * floor(dv[c] / Lb[c] + 0.5);
    }
    const double dist = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    const double mag = k * (dist - len);
    for (int c = 0; c < 3; ++c) {
      const double f = -mag * dv[c] / dist;
      F[c] = second ? F[c] - f : F[c] + f;
    }
  }
  for (int c = 0; c < 3; ++c) {
    S.F[3 * g + c] = F[c];
    S.Fa[3 * g + c] = F[c];
  }
}

// ---------------------------------------------------------------------
// residual (ErrorCheckFCM, MobilityConstraints.cpp:447-770)
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_residual(SysDev S, int nt, const double *__restrict__ V,
               const double *__restrict__ W, double *__restrict__ err,
               unsigned long long *__restrict__ maxbits) {
  if (S.act && !*S.act) return;
  if (S.ntp) nt = *S.ntp;
  __shared__ double s_m[256];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  double mx = 0.0;
  bool nonfinite = false;
  if (g < S.np) {
    const int f = S.sfil[g], i = S.sseg[g];
    const int nw = S.nw[f];
    const size_t k0 = 6 * (size_t)S.boff[f];
    const double dt = S.dt;
    const bool euler = nt <= S.impl_euler_steps;
    const double fourthirds = 4. / 3., onethird = 1. / 3.,
                 twothirdsDt = dt * 2. / 3.;
    double dw[3];
    d_dexpinv(S.Us + 3 * g, W + 3 * g, dw);
    for (int d = 0; d < 3; ++d) {
      double e1, e2;
      if (euler) {
        e1 = S.Xs[3 * g + d] - S.X[3 * g + d] - dt * V[3 * g + d];
        e2 = S.Us[3 * g + d] - dt * dw[d];
      } else {
        e1 = S.Xs[3 * g + d] - fourthirds * S.X[3 * g + d] +
             onethird * S.Xt[3 * g + d] - twothirdsDt * V[3 * g + d];
        e2 = S.Us[3 * g + d] - onethird * S.U[3 * g + d] - twothirdsDt * dw[d];
      }
      err[k0 + 3 * i + d] = e1;
      err[k0 + 3 * nw + 3 * i + d] = e2;
      // the reference's test |e| > TOL never counts a NaN, and fmax drops
      // it the same way; non-finite entries are reported on their own
      mx = fmax(mx, fmax(fabs(e1), fabs(e2)));
      nonfinite = nonfinite || !(fabs(e1) <= DBL_MAX) || !(fabs(e2) <= DBL_MAX);
    }
  }
  s_m[threadIdx.x] = mx;
  if (__syncthreads_or(nonfinite) && threadIdx.x == 0) atomicOr(maxbits + 3, 1ull);
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      s_m[threadIdx.x] = fmax(s_m[threadIdx.x], s_m[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicMax(maxbits, (unsigned long long)__double_as_longlong(s_m[0]));
}

// clears the residual's maximum (word 0) and non-finite flag (word 3)
__global__ void k_resid_reset(const int *__restrict__ act,
                              unsigned long long *mb) {
  if (act && !*act) return;
  mb[0] = 0ull;
  mb[3] = 0ull;
}

__global__ void k_fill_i32(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------
// Jacobian (Filament.hpp:803-1120)
// ---------------------------------------------------------------------
struct M3 {
  double a[3][3];
};
__device__ __forceinline__ M3 rcross(const double *x) {
  M3 A;
  A.a[0][0] = 0.;
  A.a[0][1] = x[2];
  A.a[0][2] = -x[1];
  A.a[1][0] = -x[2];
  A.a[1][1] = 0.;
  A.a[1][2] = x[0];
  A.a[2][0] = x[1];
  A.a[2][1] = -x[0];
  A.a[2][2] = 0.;
  return A;
}
__device__ __forceinline__ M3 mm(const M3 &A, const M3 &B) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j] +
                  A.a[i][2] * B.a[2][j];
  return C;
}
__device__ __forceinline__ M3 tr(const M3 &A) {
  M3 B;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B.a[i][j] = A.a[j][i];
  return B;
}
__device__ __forceinline__ M3 outer(const double *u, const double *v) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.a[i][j] = u[i] * v[j];
  return C;
}
__device__ __forceinline__ void put3(double *J, int n, int r, int c,
                                     const M3 &B) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[(size_t)(r + i) * n + c + j] = B.a[i][j];
}
__device__ __forceinline__ M3 diag_block(double dt, const M3 &dO, const M3 &LM,
                                         const M3 &OM, const double *Us) {
  M3 LdO = mm(LM, dO), LLdO = mm(LM, LdO), LOM = mm(LM, OM);
  double w[3];
  for (int i = 0; i < 3; ++i)
    w[i] = OM.a[i][0] * Us[0] + OM.a[i][1] * Us[1] + OM.a[i][2] * Us[2];
  M3 R = rcross(w), out;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out.a[i][j] = -2 * dt *
                    (dO.a[i][j] - 0.5 * (OM.a[i][j] + LdO.a[i][j]) +
                     (R.a[i][j] + LOM.a[i][j] + LLdO.a[i][j]) / 12.) /
                    3.;
  return out;
}
__device__ __forceinline__ M3 off_block(double dt, const M3 &dO, const M3 &LM) {
  M3 LdO = mm(LM, dO), LLdO = mm(LM, LdO), out;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out.a[i][j] =
          -2 * dt * (dO.a[i][j] - 0.5 * LdO.a[i][j] + LLdO.a[i][j] / 12.) / 3.;
  return out;
}

struct Frame {
  double t[3], n[3], b[3];
};
__device__ __forceinline__ Frame get_frame(const double *q) {
  Frame F;
  q_tangent(q, F.t);
  q_normal(q, F.n);
  q_binormal(q, F.b);
  return F;
}

// BlockE rows of bead k (diagonal and neighbour blocks).
__device__ void jac_blockE(const SysDev &S, double *J, int n, int f, int k,
                           int nw, int b0, int l0) {
  const fs_fil_props pr = S.props[f];
  const double dt = S.dt, DL = S.dL, MU = S.mu;
  const double myC = pr.c, myKAP = pr.kap;
  const double *ST = S.st + 3 * f;
  const int R0 = 3 * nw, C0 = 3;
  const double Tfac = 1 / (8 * kPi * MU * pow(pr.radius, 3.));
  M3 LM = tr(rcross(S.Us + 3 * (b0 + k)));
  double tau[3];
  for (int d = 0; d < 3; ++d) tau[d] = Tfac * S.TAU[3 * (b0 + k) + d];
  M3 OM = rcross(tau);
  Frame f0 = get_frame(S.qs + 4 * (b0 + k));
  M3 rt = rcross(f0.t), rn = rcross(f0.n), rb = rcross(f0.b);
  M3 EP, dO, Blk;
  if (k == 0) {
    Frame fp1 = get_frame(S.qs + 4 * (b0 + 1));
    M3 rl = rcross(S.lam + 3 * (l0 + 0));
    M3 CP = mm(rl, rt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) CP.a[i][j] = -0.5 * DL * CP.a[i][j];
    double beta = -ST[2] + 0.5 * (ddot(fp1.n, f0.b) - ddot(f0.n, fp1.b)) / DL;
    double c1[3], c2[3], cp[3], ts[3];
    dcross(f0.b, fp1.n, c1);
    dcross(f0.n, fp1.b, c2);
    for (int d = 0; d < 3; ++d) {
      cp[d] = c1[d] - c2[d];
      ts[d] = 0.5 * (f0.t[d] + fp1.t[d]);
    }
    M3 tmp1 = outer(ts, cp);
    M3 rtp1 = rcross(fp1.t);
    M3 tmp2 = mm(rtp1, rt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        EP.a[i][j] = 0.5 * myC * (tmp1.a[i][j] / DL + beta * rt.a[i][j]);
        EP.a[i][j] += 0.5 * myKAP *
                      (tmp2.a[i][j] / DL -
                       0.5 * (ST[0] * rn.a[i][j] + ST[1] * rb.a[i][j]));
        dO.a[i][j] = Tfac * (EP.a[i][j] + CP.a[i][j]);
      }
    Blk = diag_block(dt, dO, LM, OM, S.Us + 3 * (b0 + k));
    put3(J, n, R0, C0, Blk);
    M3 rnp1 = rcross(fp1.n), rbp1 = rcross(fp1.b);
    tmp2 = mm(tr(rt), rtp1);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        EP.a[i][j] = 0.5 * myC * (beta * rtp1.a[i][j] - tmp1.a[i][j] / DL);
        EP.a[i][j] += 0.5 * myKAP *
                      (tmp2.a[i][j] / DL -
                       0.5 * (ST[0] * rnp1.a[i][j] + ST[1] * rbp1.a[i][j]));
        dO.a[i][j] = Tfac * EP.a[i][j];
      }
    Blk = off_block(dt, dO, LM);
    put3(J, n, R0, C0 + 3, Blk);
  } else if (k == nw - 1) {
    Frame fm1 = get_frame(S.qs + 4 * (b0 + k - 1));
    M3 rl = rcross(S.lam + 3 * (l0 + nw - 2));
    M3 CP = mm(rl, rt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) CP.a[i][j] = -0.5 * DL * CP.a[i][j];
    double beta = -ST[2] + 0.5 * (ddot(f0.n, fm1.b) - ddot(fm1.n, f0.b)) / DL;
    double c1[3], c2[3], cp[3], ts[3];
    dcross(fm1.b, f0.n, c1);
    dcross(fm1.n, f0.b, c2);
    for (int d = 0; d < 3; ++d) {
      cp[d] = c1[d] - c2[d];
      ts[d] = 0.5 * (f0.t[d] + fm1.t[d]);
    }
    M3 tmp1 = outer(ts, cp);
    M3 rtm1 = rcross(fm1.t);
    M3 tmp2 = mm(tr(rtm1), rt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        EP.a[i][j] = -0.5 * myC * (beta * rt.a[i][j] - tmp1.a[i][j] / DL);
        EP.a[i][j] -= 0.5 * myKAP *
                      (tmp2.a[i][j] / DL -
                       0.5 * (ST[0] * rn.a[i][j] + ST[1] * rb.a[i][j]));
        dO.a[i][j] = Tfac * (EP.a[i][j] + CP.a[i][j]);
      }
    Blk = diag_block(dt, dO, LM, OM, S.Us + 3 * (b0 + k));
    put3(J, n, R0 + 3 * k, C0 + 3 * k, Blk);
    M3 rnm1 = rcross(fm1.n), rbm1 = rcross(fm1.b);
    tmp2 = mm(rt, rtm1);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        EP.a[i][j] = -0.5 * myC * (beta * rtm1.a[i][j] - tmp1.a[i][j] / DL);
        EP.a[i][j] -= 0.5 * myKAP *
                      (tmp2.a[i][j] / DL -
                       0.5 * (ST[0] * rnm1.a[i][j] + ST[1] * rbm1.a[i][j]));
        dO.a[i][j] = Tfac * EP.a[i][j];
      }
    Blk = off_block(dt, dO, LM);
    put3(J, n, R0 + 3 * k, C0 + 3 * k - 3, Blk);
  } else {
    Frame fm1 = get_frame(S.qs + 4 * (b0 + k - 1));
    Frame fp1 = get_frame(S.qs + 4 * (b0 + k + 1));
    double lv[3];
    for (int d = 0; d < 3; ++d)
      lv[d] = S.lam[3 * (l0 + k - 1) + d] + S.lam[3 * (l0 + k) + d];
    M3 rl = rcross(lv);
    M3 CP = mm(rl, rt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) CP.a[i][j] = -0.5 * DL * CP.a[i][j];
    double beta = -ST[2] + 0.5 * (ddot(fp1.n, f0.b) - ddot(f0.n, fp1.b)) / DL;
    double c1[3], c2[3], cp[3], ts[3], tsm[3];
    dcross(f0.b, fp1.n, c1);
    dcross(f0.n, fp1.b, c2);
    for (int d = 0; d < 3; ++d) {
      cp[d] = c1[d] - c2[d];
      ts[d] = 0.5 * (f0.t[d] + fp1.t[d]);
      tsm[d] = 0.5 * (f0.t[d] + fm1.t[d]);
    }
    M3 tmp1 = outer(ts, cp), tmp3 = outer(tsm, cp);
    M3 rtp1 = rcross(fp1.t), rtm1 = rcross(fm1.t);
    M3 tmp2 = mm(rtp1, rt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        EP.a[i][j] = 0.5 * myC * (tmp1.a[i][j] / DL + beta * rt.a[i][j]);
        EP.a[i][j] += 0.5 * myKAP *
                      (tmp2.a[i][j] / DL -
                       0.5 * (ST[0] * rn.a[i][j] + ST[1] * rb.a[i][j]));
      }
    // backward side: beta recomputed, cross part reused (Filament.hpp:1065)
    beta = -ST[2] + 0.5 * (ddot(f0.n, fm1.b) - ddot(fm1.n, f0.b)) / DL;
    tmp2 = mm(tr(rtm1), rt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        EP.a[i][j] -= 0.5 * myC * (beta * rt.a[i][j] - tmp3.a[i][j] / DL);
        EP.a[i][j] -= 0.5 * myKAP *
                      (tmp2.a[i][j] / DL -
                       0.5 * (ST[0] * rn.a[i][j] + ST[1] * rb.a[i][j]));
        dO.a[i][j] = Tfac * (CP.a[i][j] + EP.a[i][j]);
      }
    Blk = diag_block(dt, dO, LM, OM, S.Us + 3 * (b0 + k));
    put3(J, n, R0 + 3 * k, C0 + 3 * k, Blk);
    M3 rnm1 = rcross(fm1.n), rbm1 = rcross(fm1.b);
    tmp2 = mm(rt, rtm1);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        EP.a[i][j] = -0.5 * myC * (beta * rtm1.a[i][j] - tmp3.a[i][j] / DL);
        EP.a[i][j] -= 0.5 * myKAP *
                      (tmp2.a[i][j] / DL -
                       0.5 * (ST[0] * rnm1.a[i][j] + ST[1] * rbm1.a[i][j]));
        dO.a[i][j] = Tfac * EP.a[i][j];
      }
    Blk = off_block(dt, dO, LM);
    put3(J, n, R0 + 3 * k, C0 + 3 * k - 3, Blk);
    M3 rnp1 = rcross(fp1.n), rbp1 = rcross(fp1.b);
    tmp2 = mm(tr(rt), rtp1);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        EP.a[i][j] = 0.5 * myC * (beta * rtp1.a[i][j] - tmp1.a[i][j] / DL);
        EP.a[i][j] += 0.5 * myKAP *
                      (tmp2.a[i][j] / DL -
                       0.5 * (ST[0] * rnp1.a[i][j] + ST[1] * rbp1.a[i][j]));
        dO.a[i][j] = Tfac * EP.a[i][j];
      }
    Blk = off_block(dt, dO, LM);
    put3(J, n, R0 + 3 * k, C0 + 3 * k + 3, Blk);
  }
  for (int d = 0; d < 3; ++d)
    J[(size_t)(R0 + 3 * k + d) * n + C0 + 3 * k + d] += 1.;
}

__global__ void k_fill_f64(double *p, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// buildApproximateJacobian (Filament.hpp:803-867), one workgroup per
// filament, one lane per bead / link.
__global__ void k_jac_build(SysDev S, int nt, double *__restrict__ Jall) {
  if (S.ntp) nt = *S.ntp;
  const int f = blockIdx.x;
  const int b0 = S.boff[f], l0 = S.loff[f], nw = S.nw[f];
  const int n = 6 * nw;
  double *J = Jall + S.joff[f];
  const fs_fil_props pr = S.props[f];
  const double dt = S.dt, DL = S.dL, MU = S.mu;
  const double rad = pr.radius;
  if (threadIdx.x == 0) {
    int jj = 3 * nw + 3;
    double temp = dt / (9 * kPi * MU * rad);
    for (int d = 0; d < 3; ++d) {
      J[(size_t)d * n + d] = 1;
      J[(size_t)d * n + jj + d] = temp / dt;
    }
  }
  const double Jprefac = 10.;
  // BlockC(0.5 DL): row block r (bead r+1) = [f R(t0), 2f R(t1) .. 2f R(tr),
  // f R(t_{r+1}), 0 ..]: the reference's cumulative copy (Filament.hpp:
  // 885-909) sums each shared term twice, exactly.
  const double fac = 0.5 * DL;
  for (int r = threadIdx.x; r < nw - 1; r += blockDim.x) {
    for (int c = 0; c <= r + 1; ++c) {
      double t[3];
      q_tangent(S.qs + 4 * (b0 + c), t);
      M3 Rt = rcross(t);
      M3 B;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double v = fac * Rt.a[i][j];
          B.a[i][j] = (c >= 1 && c <= r) ? v + v : v;
        }
      put3(J, n, 3 + 3 * r, 3 + 3 * c, B);
    }
    // Jprefac * BlockD(nt) / dt (Filament.hpp:913-946)
    int m = r;
    double facn = -2. * dt / (18. * kPi * MU * rad);
    if (nt <= S.impl_euler_steps) facn *= (3. / 2.);
    int i = 3 * m;
    for (int d = 0; d < 3; ++d)
      J[(size_t)(3 + i + d) * n + 3 * nw + 3 + i + d] = Jprefac * facn / dt;
    if (m < nw - 2) {
      facn = 2. * dt / (18. * kPi * MU * rad);
      if (nt <= S.impl_euler_steps) facn *= (3. / 2.);
      for (int d = 0; d < 3; ++d)
        J[(size_t)(3 + i + d) * n + 3 * nw + 3 + i + 3 + d] = Jprefac * facn / dt;
    }
    // BlockF / dt (Filament.hpp:1098-1120): link r couples beads r, r+1
    for (int side = 0; side < 2; ++side) {
      int bead = r + side;
      double fc = (side == 0) ? dt * DL / (24 * kPi * MU * pow(rad, 3.))
                              : dt * DL / (24. * kPi * MU * pow(rad, 3));
      double tv[3];
      q_tangent(S.qs + 4 * (b0 + bead), tv);
      M3 tmat = tr(rcross(tv));
      M3 umat = rcross(S.Us + 3 * (b0 + bead));
      M3 um = mm(umat, tmat);
      M3 uu = mm(umat, um);
      M3 B;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          B.a[a][b] = fc * (tmat.a[a][b] - 0.5 * um.a[a][b] + uu.a[a][b] / 12.) / dt;
      put3(J, n, 3 * nw + 3 * r + 3 * side, 3 * nw + 3 + 3 * r, B);
    }
  }
  for (int k = threadIdx.x; k < nw; k += blockDim.x)
    jac_blockE(S, J, n, f, k, nw, b0, l0);
}

// In-place Gauss-Jordan with partial pivoting (pivot rows are not moved:
// the step that pivots on row p_k is unscrambled at the end as
// A^-1[k][p_m] = M[p_k][m]). Output written column-major (Jinv^T) with the
// lambda rows scaled by 1/dt.
__global__ void __launch_bounds__(1024)
    k_gj_inverse(SysDev S, double *__restrict__ Jall,
                 double *__restrict__ JTall, int *__restrict__ singular) {
  __shared__ double colbuf[6 * kMaxNw / 4 * 4];
  __shared__ double rowbuf[6 * kMaxNw / 4 * 4];
  __shared__ int used[6 * kMaxNw / 4 * 4], pivrow[6 * kMaxNw / 4 * 4],
      rowstep[6 * kMaxNw / 4 * 4];
  __shared__ int s_p;
  const int f = blockIdx.x;
  const int nw = S.nw[f];
  const int n = 6 * nw;
  double *M = Jall + S.joff[f];
  double *JT = JTall + S.joff[f];
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += blockDim.x) used[i] = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    for (int i = tid; i < n; i += blockDim.x) colbuf[i] = M[(size_t)i * n + k];
    __syncthreads();
    if (tid < 64) {
      // used rows rank -1 and unused NaN entries -0.5, so the pivot is
      // always an unused row (n - k of them remain) and stays in range; a
      // pivot that is not > 0 or not finite raises the singular flag
      double best = -2.0;
      int bi = n;
      for (int i = tid; i < n; i += 64) {
        const double a = fabs(colbuf[i]);
        const double v = used[i] ? -1.0 : (a == a ? a : -0.5);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        double ob = __shfl_xor(best, off, 64);
        int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (tid == 0) {
        if (!(best > 0.0) || !(best <= DBL_MAX) || bi >= n)
          atomicExch(singular, 1);
        bi = bi < n ? bi : n - 1;
        s_p = bi;
        used[bi] = 1;
        pivrow[k] = bi;
        rowstep[bi] = k;
      }
    }
    __syncthreads();
    const int p = s_p;
    for (int j = tid; j < n; j += blockDim.x) rowbuf[j] = M[(size_t)p * n + j];
    __syncthreads();
    const double inv = 1.0 / rowbuf[k];
    for (int e = tid; e < n * n; e += blockDim.x) {
      int i = e / n, j = e - i * n;
      if (i == p) {
        M[e] = (j == k) ? inv : rowbuf[j] * inv;
      } else {
        double fct = colbuf[i];
        if (j == k)
          M[e] = -fct * inv;
        else if (fct != 0.0)
          M[e] = M[e] - fct * (rowbuf[j] * inv);
      }
    }
    __syncthreads();
  }
  const double idt = 1.0 / S.dt;
  for (int e = tid; e < n * n; e += blockDim.x) {
    int i = e / n, j = e - i * n;
    int r = rowstep[i], c = pivrow[j];
    double v = M[e];
    if (r >= 3 * nw + 3) v /= S.dt;
    JT[(size_t)c * n + r] = v;
  }
  (void)idt;
}

// Block-tridiagonal factorisation of the approximate Jacobian.
// Differencing the position rows of beads k >= 2 (row k minus row k - 1)
// turns the cumulative BlockC rows (Filament.hpp:885-909) into two blocks
// and BlockD into three; ordering the unknowns per bead, [X0 U_0 lam_0 |
// U_1 lam_1 | .. | U_{Nw-1}], and the equations per bead, [P_k L_k] (P_0
// the first-bead rows), leaves A^ = P_r T J P_c^T with at most 8 sub- and 8
// superdiagonals (the zeros outside are exact: the cumulative entries of
// neighbouring rows are equal). Cut into blocks of kBlk = 12 rows and
// columns, A^ is block tridiagonal, and
//   J^-1 b = P_c^T A^^-1 P_r T b
// follows from the block LU recursion (block Thomas):
//   D_K = A_KK - A_K,K-1 G_K-1,  G_K = D_K^-1 A_K,K+1,
//   y_K = D_K^-1 (b_K - A_K,K-1 y_K-1),  x_K = y_K - G_K x_K+1,
// with each D_K inverted by Gauss-Jordan with partial pivoting inside the
// block. That replaces the dense 6Nw x 6Nw inverse (arma::inv,
// Filament.hpp:862): O(Nw) work and storage, and a solve reads 3 small
// blocks per 12 rows instead of a dense row of J^-1 per row. (The solution
// agrees with the dense inverse to rounding on the reference's Jacobians,
// including stiff swimmers; a block D_K that is singular or not finite
// raises the singular flag.)
constexpr int kBlk = 12;
constexpr int kBlkSz = kBlk * kBlk;
constexpr int kBlkRec = 3 * kBlkSz;  // D^-1, G, A_K,K-1 per block row
// Lane permutations of the data-parallel primitives (DPP): a few cycles,
// against a round trip through the LDS crossbar for a general shuffle.
// Quad permutations [1,0,3,2] and [2,3,0,1] (lane l ^ 1, l ^ 2), and the
// mirrors inside each group of 8 and of 16 lanes.
constexpr int kDppXor1 = 1 | 0 << 2 | 3 << 4 | 2 << 6;
constexpr int kDppXor2 = 2 | 3 << 2 | 0 << 4 | 1 << 6;
constexpr int kDppMirror = 0x140, kDppHalfMirror = 0x141;
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double a) {
  const long long v = __double_as_longlong(a);
  const int lo = dpp_i32<CTRL>((int)v);
  const int hi = dpp_i32<CTRL>((int)(v >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// a's value in lane l ^ X (X = 1 or 2, inside each group of four lanes)
template <int X>
__device__ __forceinline__ double quad_xor(double a) {
  static_assert(X == 1 || X == 2, "quad lanes");
  return dpp_f64<X == 1 ? kDppXor1 : kDppXor2>(a);
}
constexpr int kBandMaxNwBlk = 32;  // block rows of the longest filament (Nw <= 64)

// Original unknown (column of J) of permuted unknown j, and original
// equation (row of J) of permuted equation i; *diff: the row is differenced
// with row - 3.
__device__ __forceinline__ int band_col(int j, int nw) {
  if (j < 3) return j;  // X0
  const int k = (j - 3) / 6, w = (j - 3) - 6 * k;
  return w < 3 ? 3 + 3 * k + w : 3 * nw + 3 * k + w;  // U_k, lam_k
}
__device__ __forceinline__ int band_row(int i, int nw, bool *diff) {
  const int k = i / 6, w = i - 6 * k;
  *diff = (w < 3) && k >= 2;
  return w < 3 ? 3 * k + w : 3 * nw + 3 * k + (w - 3);  // P_k, L_k
}
// A^(i, j) from the dense J (n x n row-major); 0 outside [0, n)^2
__device__ __forceinline__ double band_entry(const double *J, int n, int nw,
                                             int i, int j) {
  if (i >= n || j >= n || i < 0 || j < 0) return 0.0;
  bool diff;
  const int ro = band_row(i, nw, &diff), co = band_col(j, nw);
  double v = J[(size_t)ro * n + co];
  if (diff) v -= J[(size_t)(ro - 3) * n + co];
  return v;
}
// Block factors of filament f start here (block rows <= Nw / 2 + 1 each).
__device__ __forceinline__ size_t blk_offset(const SysDev &S, int f) {
  return ((size_t)S.boff[f] + 2 * (size_t)f) * (kBlkRec / 2);
}

// One wavefront per filament (a one-wave workgroup: its barriers are
// cheap, and the pivot search needs no second wave): block rows in order,
// the 12 x 12 operations spread over the lanes (padding rows past n carry
// an identity diagonal).
constexpr int kFacT = 64;
__global__ void __launch_bounds__(kFacT)
    k_blk_factor(SysDev S, const double *__restrict__ Jall,
                 double *__restrict__ Fall, int *__restrict__ singular) {
  if (S.act && !*S.act) return;
  __shared__ double sA[kBlk][2 * kBlk];  // [D | I] -> [I | D^-1]
  __shared__ double sL[kBlkSz], sU[kBlkSz], sG[kBlkSz];
  const int f = blockIdx.x;
  const int nw = S.nw[f], n = 6 * nw, nb = (n + kBlk - 1) / kBlk;
  const double *J = Jall + S.joff[f];
  double *F = Fall + blk_offset(S, f);
  const int tid = threadIdx.x;
  bool bad = false;
  for (int e = tid; e < kBlkSz; e += kFacT) sG[e] = 0.0;
  for (int K = 0; K < nb; ++K) {
    const int r0 = K * kBlk;
    // A_K,K-1 and A_K,K+1 of this block row; D = A_KK (identity padding)
    for (int e = tid; e < kBlkSz; e += kFacT) {
      const int i = e / kBlk, j = e - i * kBlk;
      sL[e] = band_entry(J, n, nw, r0 + i, r0 - kBlk + j);
      sU[e] = band_entry(J, n, nw, r0 + i, r0 + kBlk + j);
      double d = band_entry(J, n, nw, r0 + i, r0 + j);
      if (r0 + i >= n && i == j) d = 1.0;
      sA[i][j] = d;
      sA[i][kBlk + j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    // D -= A_K,K-1 G_K-1
    if (K > 0)
      for (int e = tid; e < kBlkSz; e += kFacT) {
        const int i = e / kBlk, j = e - i * kBlk;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < kBlk; ++m) acc += sL[i * kBlk + m] * sG[m * kBlk + j];
        sA[i][j] -= acc;
      }
    __syncthreads();
    // [D | I] -> [I | D^-1], partial pivoting over the rows below, in
    // registers: lane 4 r + q (r < 12) holds columns 6q .. 6q + 5 of row r
    {
      const int gr = tid >> 2, gq = tid & 3;
      const bool own = tid < 4 * kBlk;
      double a[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) a[k] = own ? sA[gr][6 * gq + k] : 0.0;
#pragma unroll
      for (int c = 0; c < kBlk; ++c) {
        const int cq = c / 6, ck = c % 6;
        // pivot: the largest |A[r][c]| over r >= c (NaN below every number,
        // ties to the lower row), compared as bit patterns of the absolute
        // values in uniform registers
        unsigned long long best = 0, wbits = 0;
        int p = c;
#pragma unroll
        for (int r = c; r < kBlk; ++r) {
          const long long v = __double_as_longlong(a[ck]);
          const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)v, 4 * r + cq);
          const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), 4 * r + cq);
          const unsigned long long bits =
              (((unsigned long long)hi << 32) | lo) & 0x7FFFFFFFFFFFFFFFull;
          const unsigned long long key =
              bits > 0x7FF0000000000000ull ? 1ull : bits + 2ull;
          if (key > best) {
            best = key;
            wbits = bits;
            p = r;
          }
        }
        // zero, NaN or infinite pivot: singular (the row order still holds)
        bad = bad || best < 2 || wbits == 0 || wbits >= 0x7FF0000000000000ull;
        if (p != c) {  // uniform: rows c and p trade their registers
          const int d = 4 * (p - c);
          const int src = gr == c ? tid + d : (gr == p ? tid - d : tid);
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            const long long v = __double_as_longlong(a[k]);
            const int lo = __shfl((int)v, src, 64);
            const int hi = __shfl((int)(v >> 32), src, 64);
            a[k] = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
          }
        }
        double pv;
        {
          const long long v = __double_as_longlong(a[ck]);
          const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)v, 4 * c + cq);
          const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), 4 * c + cq);
          pv = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
        }
        const double inv = 1.0 / pv;
        // eliminate column c from the other rows: A[r][j] -= A[r][c] inv A[c][j]
        // (A[r][c] from the lane of this row holding column c, A[c][j] from
        // the lane of row c holding column j), then scale row c
        const double arc = cq == 0 ? dpp_f64<0x00>(a[ck])
                                   : (cq == 1 ? dpp_f64<0x55>(a[ck])
                                              : (cq == 2 ? dpp_f64<0xAA>(a[ck])
                                                         : dpp_f64<0xFF>(a[ck])));
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const long long v = __double_as_longlong(a[k]);
          const int lo = __shfl((int)v, 4 * c + gq, 64);
          const int hi = __shfl((int)(v >> 32), 4 * c + gq, 64);
          const double acj = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
          a[k] = gr == c ? a[k] * inv : a[k] - arc * inv * acj;
        }
      }
      if (own)
#pragma unroll
        for (int k = 0; k < 6; ++k) sA[gr][6 * gq + k] = a[k];
      __syncthreads();
    }
    // G_K = D^-1 A_K,K+1; store D^-1, G_K, A_K,K-1
    double *FK = F + (size_t)K * kBlkRec;
    for (int e = tid; e < kBlkSz; e += kFacT) {
      const int i = e / kBlk, j = e - i * kBlk;
      double acc = 0.0;
#pragma unroll
      for (int m = 0; m < kBlk; ++m) acc += sA[i][kBlk + m] * sU[m * kBlk + j];
      sG[e] = acc;
      FK[e] = sA[i][kBlk + j];
      FK[kBlkSz + e] = acc;
      FK[2 * kBlkSz + e] = sL[e];
    }
    __syncthreads();
  }
  if (tid == 0 && bad) atomicExch(singular, 1);
}

// x = J^-1 b for every filament (FilamentJacobianSolver.hpp:49-73 without
// the history), lambda rows scaled by 1/dt as the reference's inverse
// (Filament.hpp:863-865). One wavefront per filament, forward then backward
// block sweeps: lane 4 row + part (row < 12) sums 3 of the 12 products of
// a row of a 12 x 12 factor, the four parts combine by lane exchange; the
// right-hand side lives in LDS, each lane's factor entries in registers,
// loaded kSolveAhead block rows ahead of their use (a sweep step is much
// shorter than a global load's latency).
constexpr int kSolveAhead = 6;
// Ordering of LDS accesses between the lanes of a one-wavefront workgroup:
// a wavefront-scope fence pair around a wave barrier. Unlike a workgroup
// barrier it does not wait for the wave's outstanding global loads, so
// loads issued rows ahead stay in flight across it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__global__ void __launch_bounds__(64)
    k_blk_solve(SysDev S, const double *__restrict__ Fall,
                const double *__restrict__ b, double *__restrict__ x) {
  if (S.act && !*S.act) return;
  __shared__ double sy[kBandMaxNwBlk * kBlk];  // permuted b, then y, x
  __shared__ double sr[kBlk];
  const int f = blockIdx.x;
  const int nw = S.nw[f], n = 6 * nw, nb = (n + kBlk - 1) / kBlk;
  const size_t so = 6 * (size_t)S.boff[f];
  const double *F = Fall + blk_offset(S, f);
  const int lane = threadIdx.x;
  for (int i = lane; i < nb * kBlk; i += 64) {
    double y = 0.0;
    if (i < n) {
      bool diff;
      const int ro = band_row(i, nw, &diff);
      y = b[so + ro];
      if (diff) y -= b[so + ro - 3];
    }
    sy[i] = y;
  }
  const int row = lane >> 2, part = lane & 3;
  const bool act = row < kBlk;
  const int mo = act ? row * kBlk + 3 * part : 0;  // this lane's 3 entries
  // every lane loads (lanes past row 11 read row 0's entries and zero
  // them), so no load sits behind a branch and the waits before a slot's
  // use count only the loads issued after it
  auto load3 = [&](const double *M, double *o) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const double v = M[mo + m];
      o[m] = act ? v : 0.0;
    }
  };
  auto dot3 = [&](const double *o, const double *v) {
    double a = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) a += o[m] * v[3 * part + m];
    a += quad_xor<1>(a);
    a += quad_xor<2>(a);
    return a;
  };
  // forward: y_K = D_K^-1 (b_K - A_K,K-1 y_K-1). Register slot j holds the
  // factor entries of the rows K = j (mod kSolveAhead): each row's slot is
  // refilled, right after its use, with the row kSolveAhead further on (the
  // last row again past the end: every load is issued, so the wait before
  // a slot's use counts only the loads issued after it)
  constexpr int P = kSolveAhead;
  double dq[P][3], lq[P][3];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int r = min(j, nb - 1);
    load3(F + (size_t)r * kBlkRec, dq[j]);
    load3(F + (size_t)r * kBlkRec + 2 * kBlkSz, lq[j]);
  }
  wave_sync();
  for (int K0 = 0; K0 < nb; K0 += P) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int K = K0 + j;
      if (K < nb) {
        double *yK = sy + K * kBlk;
        const double t = K > 0 ? dot3(lq[j], yK - kBlk) : 0.0;
        if (act && part == 0) sr[row] = yK[row] - t;
        wave_sync();
        const double y = dot3(dq[j], sr);
        if (act && part == 0) yK[row] = y;
        wave_sync();
      }
      const int r = min(K + P, nb - 1);
      load3(F + (size_t)r * kBlkRec, dq[j]);
      load3(F + (size_t)r * kBlkRec + 2 * kBlkSz, lq[j]);
    }
  }
  // backward: x_K = y_K - G_K x_K+1 over K = nb - 2 .. 0, the same slots
  // for the G of rows nb - 2 - j (mod kSolveAhead)
  double gq[P][3];
#pragma unroll
  for (int j = 0; j < P; ++j)
    load3(F + (size_t)max(nb - 2 - j, 0) * kBlkRec + kBlkSz, gq[j]);
  for (int s0 = 0; s0 < nb - 1; s0 += P) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int K = nb - 2 - (s0 + j);
      if (K >= 0) {
        double *yK = sy + K * kBlk;
        const double t = dot3(gq[j], yK + kBlk);
        if (act && part == 0) yK[row] -= t;
        wave_sync();
      }
      load3(F + (size_t)max(K - P, 0) * kBlkRec + kBlkSz, gq[j]);
    }
  }
  const double idt = 1.0 / S.dt;
  for (int i = lane; i < n; i += 64) {
    const int co = band_col(i, nw);
    x[so + co] = co >= 3 * nw + 3 ? sy[i] * idt : sy[i];
  }
}

// ---------------------------------------------------------------------
// Broyden solve (FilamentJacobianSolver.hpp:49-73)
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_jsolve_block(SysDev S, const double *__restrict__ JTall,
                   const double *__restrict__ b, double *__restrict__ x) {
  if (S.act && !*S.act) return;
  __shared__ double sb[6 * kMaxNw];
  const int f = blockIdx.x;
  const int nw = S.nw[f], n = 6 * nw;
  const size_t so = 6 * (size_t)S.boff[f];
  const double *JT = JTall + S.joff[f];
  for (int i = threadIdx.x; i < n; i += blockDim.x) sb[i] = b[so + i];
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    double s = 0;
    for (int c = 0; c < n; ++c) s += JT[(size_t)c * n + r] * sb[c];
    x[so + r] = s;
  }
}

// partial[blk * ldp + i] = sum over rows of chunk blk of A_i[r] * b[r]
// (A column-major, column stride lda) for i < ncol. Fixed partition:
// deterministic. With ncolp the column count is *ncolp + noff, read on the
// device (batched sweep: the grid covers the largest count, blocks past it
// return); act: skip the launch when *act == 0.
__global__ void __launch_bounds__(256)
    k_dots_partial(const int *__restrict__ act, const double *__restrict__ A,
                   size_t lda, int ncol, const int *__restrict__ ncolp,
                   int noff, int ldp, const double *__restrict__ b,
                   size_t nrow, int chunk, double *__restrict__ partial) {
  if (act && !*act) return;
  if (ncolp) ncol = *ncolp + noff;
  __shared__ double red[256];
  const int blk = blockIdx.x, i = blockIdx.y;
  if (i >= ncol) return;
  const size_t lo = (size_t)blk * chunk;
  const size_t hi = min(nrow, lo + chunk);
  const double *Ai = A + lda * i;
  double s = 0;
  for (size_t r = lo + threadIdx.x; r < hi; r += 256) s += Ai[r] * b[r];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)blk * ldp + i] = red[0];
}

// one 64-lane workgroup per column: strided lane sums, then a butterfly in
// a fixed order (deterministic)
__global__ void __launch_bounds__(64)
    k_dots_final(const int *__restrict__ act, const double *__restrict__ partial,
                 int nblk, int ldp, int ncol, const int *__restrict__ ncolp,
                 int noff, double *__restrict__ out) {
  if (act && !*act) return;
  if (ncolp) ncol = *ncolp + noff;
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= ncol) return;
  double s = 0;
  for (int b = lane; b < nblk; b += 64) s += partial[(size_t)b * ldp + i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[i] = s;
}

__global__ void k_hist_axpy(const int *__restrict__ act,
                            const double *__restrict__ C, size_t lda, int ncol,
                            const int *__restrict__ ncolp, int noff,
                            const double *__restrict__ coef, size_t nrow,
                            double *__restrict__ x) {
  if (act && !*act) return;
  if (ncolp) ncol = *ncolp + noff;
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  double v = x[r];
  for (int i = 0; i < ncol; ++i) v += C[lda * i + r] * coef[i];
  x[r] = v;
}

__global__ void k_sub(const double *__restrict__ a, const double *__restrict__ b,
                      double *__restrict__ c, size_t n) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) c[r] = a[r] - b[r];
}

// bad-Broyden columns (main.cpp:640-663)
__global__ void k_broyden_cols(const double *__restrict__ tmp,
                               const double *__restrict__ upd,
                               const double *__restrict__ dE,
                               const double *__restrict__ nrm2, size_t n,
                               double *__restrict__ Ccol,
                               double *__restrict__ Dcol) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double nrm = sqrt(*nrm2);
  double t = tmp[r] * (-1. / nrm);
  t -= upd[r] / nrm;
  Ccol[r] = t;
  Dcol[r] = dE[r] / nrm;
}

// dE = Enew - E and the per-block partial sums of dE.dE and dE.Enew
// (fixed partition: deterministic).
__global__ void __launch_bounds__(256)
    k_diff_dots(const int *__restrict__ act, const double *__restrict__ En,
                const double *__restrict__ Eo, double *__restrict__ dE,
                size_t n, int chunk, double *__restrict__ partial) {
  if (act && !*act) return;
  __shared__ double r1[256], r2[256];
  const size_t lo = (size_t)blockIdx.x * chunk;
  const size_t hi = min(n, lo + chunk);
  double a = 0, b = 0;
  for (size_t r = lo + threadIdx.x; r < hi; r += 256) {
    double e = En[r], d = e - Eo[r];
    dE[r] = d;
    a += d * d;
    b += d * e;
  }
  r1[threadIdx.x] = a;
  r2[threadIdx.x] = b;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      r1[threadIdx.x] += r1[threadIdx.x + off];
      r2[threadIdx.x] += r2[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r1[0];
    partial[2 * blockIdx.x + 1] = r2[0];
  }
}

// Fused bad-Broyden step. With g = H_k E_new (history columns < k), the
// reference's second solve H_k dE equals g - Update, so its column
// C_k = (-H_k dE - Update)/|dE| = -g/|dE|, D_k = dE/|dE|, and the next
// update H_{k+1} E_new = g + C_k (D_k . E_new) = g (1 - dE.E_new/|dE|^2).
// Columns col of C and D (column stride n); with iterp, col = *iterp - 1.
__global__ void k_broyden_fused(const int *__restrict__ act,
                                const double *__restrict__ g,
                                const double *__restrict__ dE,
                                const double *__restrict__ dots, size_t n,
                                double *__restrict__ Cm,
                                double *__restrict__ Dm, int col,
                                const int *__restrict__ iterp,
                                double *__restrict__ upd) {
  if (act && !*act) return;
  if (iterp) col = *iterp - 1;
  double *Ccol = Cm + n * (size_t)col, *Dcol = Dm + n * (size_t)col;
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double nrm2 = dots[0], dEN = dots[1];
  const double nrm = sqrt(nrm2);
  const double gr = g[r];
  Ccol[r] = -gr / nrm;
  Dcol[r] = dE[r] / nrm;
  upd[r] = gr * (1.0 - dEN / nrm2);
}

__global__ void k_copy_f64(const int *__restrict__ act,
                           const double *__restrict__ src,
                           double *__restrict__ dst, size_t n) {
  if (act && !*act) return;
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[r] = src[r];
}

// Device control block of a suspension in a batched sweep: ctl[0] = Broyden
// iterations done this step, ctl[1] = still iterating. The loop test of
// main.cpp:494-503, 684-687 on the device: iterate while the residual's
// infinity norm exceeds TOL and fewer than broyden_maxiter iterations ran.
__global__ void k_ctl_begin(int *__restrict__ ctl,
                            const unsigned long long *__restrict__ maxbits,
                            double tol, int maxiter) {
  const double m = __longlong_as_double((long long)*maxbits);
  ctl[0] = 0;
  ctl[1] = (m > tol && maxiter > 0) ? 1 : 0;
}

__global__ void k_ctl_after(int *__restrict__ ctl,
                            const unsigned long long *__restrict__ maxbits,
                            double tol, int maxiter) {
  if (!ctl[1]) return;
  const double m = __longlong_as_double((long long)*maxbits);
  const int it = ctl[0] + 1;
  ctl[0] = it;
  ctl[1] = (m > tol && it < maxiter) ? 1 : 0;
}

// publish ints of a suspension into the batch's status array
__global__ void k_publish(const int *__restrict__ src, int n,
                          int *__restrict__ dst) {
  if (threadIdx.x < n) dst[threadIdx.x] = src[threadIdx.x];
}

// Filament::step (Filament.hpp:455-478) + Bead::step (Bead.hpp:266-276)
__global__ void k_step(SysDev S, int nt) {
  if (S.ntp) nt = *S.ntp;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (size_t)S.np) {
    size_t g = t;
    for (int d = 0; d < 3; ++d) {
      S.Xt[3 * g + d] = S.X[3 * g + d];
      S.X[3 * g + d] = S.Xs[3 * g + d];
      S.Ut[3 * g + d] = S.U[3 * g + d];
      S.U[3 * g + d] = S.Us[3 * g + d];
    }
    double q[4];
    for (int c = 0; c < 4; ++c) {
      S.qt[4 * g + c] = S.q[4 * g + c];
      q[c] = S.q[4 * g + c];
    }
    q_expm(q, S.U + 3 * g);
    for (int c = 0; c < 4; ++c) S.q[4 * g + c] = q[c];
  }
  if (t < 3 * (size_t)S.nlam) {
    size_t l = t;
    if (nt > 12) {
      double v = 3 * S.lam[l] - 3 * S.lam1[l] + S.lam2[l];
      S.lam2[l] = S.lam1[l];
      S.lam1[l] = S.lam[l];
      S.lam[l] = v;
    } else if (nt == 12) {
      double v = 2. * S.lam[l] - S.lam1[l];
      S.lam2[l] = S.lam1[l];
      S.lam1[l] = S.lam[l];
      S.lam[l] = v;
    } else {
      S.lam1[l] = S.lam[l];
    }
  }
}

// ---------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------
int SystemImpl::init(const fs_sim_params &pp, int nfil, const int *nwv,
                     const fs_fil_props *props) {
  p = pp;
  FS_CHECK(nfil >= 1, FS_ERR_ARG, "nfil must be >= 1");
  FS_CHECK(p.dt > 0 && p.dL > 0 && p.tol > 0, FS_ERR_ARG,
           "dt, dL, tol must be positive");
  FS_CHECK(p.broyden_maxiter >= 1, FS_ERR_ARG, "broyden_maxiter must be >= 1");
  nf = nfil;
  hnw.assign(nwv, nwv + nfil);
  hboff.resize(nfil + 1);
  hloff.resize(nfil + 1);
  hjoff.resize(nfil + 1);
  hboff[0] = hloff[0] = 0;
  hjoff[0] = 0;
  for (int f = 0; f < nfil; ++f) {
    FS_CHECK(nwv[f] >= 2 && nwv[f] <= kMaxNw, FS_ERR_ARG,
             "filament %d: %d segments (need 2..%d)", f, nwv[f], kMaxNw);
    hboff[f + 1] = hboff[f] + nwv[f];
    hloff[f + 1] = hloff[f] + nwv[f] - 1;
    hjoff[f + 1] = hjoff[f] + (long)(6 * nwv[f]) * (6 * nwv[f]);
  }
  np = hboff[nfil];
  nlam = hloff[nfil];
  std::vector<int> sf(np), ss(np);
  for (int f = 0; f < nfil; ++f)
    for (int i = 0; i < nwv[f]; ++i) {
      sf[hboff[f] + i] = f;
      ss[hboff[f] + i] = i;
    }
  const size_t n3 = 3 * (size_t)np, n4 = 4 * (size_t)np, l3 = 3 * (size_t)nlam;
  const size_t n6 = 6 * (size_t)np;
  // one arena for the per-bead/per-link doubles
  size_t nd = 6 * n3 + 3 * n4 + 2 * n3 + 3 * l3 + 3 * (size_t)nf + 5 * n3 +
              6 * n6;
  FS_TRY(arena.alloc(sizeof(double) * nd));
  double *a = arena.as<double>();
  auto take = [&](size_t k) {
    double *r = a;
    a += k;
    return r;
  };
  d.X = take(n3);
  d.Xs = take(n3);
  d.Xt = take(n3);
  d.U = take(n3);
  d.Us = take(n3);
  d.Ut = take(n3);
  d.q = take(n4);
  d.qs = take(n4);
  d.qt = take(n4);
  d.F = take(n3);
  d.TAU = take(n3);
  d.lam = take(l3);
  d.lam1 = take(l3);
  d.lam2 = take(l3);
  d.st = take(3 * (size_t)nf);
  d.Ya = take(n3);
  d.Fa = take(n3);
  d.Ta = take(n3);
  d.V = take(n3);
  d.W = take(n3);
  E = take(n6);
  NewE = take(n6);
  Upd = take(n6);
  tmp = take(n6);
  dE = take(n6);
  scratch6 = take(n6);
  FS_HIP(hipMemset(arena.p, 0, sizeof(double) * nd));
  // identity quaternions
  std::vector<double> qi(n4, 0.0);
  for (int g = 0; g < np; ++g) qi[4 * g] = 1.0;
  FS_HIP(hipMemcpy(d.q, qi.data(), sizeof(double) * n4, hipMemcpyHostToDevice));
  FS_HIP(hipMemcpy(d.qs, qi.data(), sizeof(double) * n4, hipMemcpyHostToDevice));
  FS_HIP(hipMemcpy(d.qt, qi.data(), sizeof(double) * n4, hipMemcpyHostToDevice));
  std::vector<double> st0(3 * (size_t)nf, 0.0);
  for (int f = 0; f < nf; ++f) st0[3 * f] = p.strain_twist0;
  FS_HIP(hipMemcpy(d.st, st0.data(), sizeof(double) * 3 * nf, hipMemcpyHostToDevice));
  FS_TRY(ints.alloc(sizeof(int) * (3 * (size_t)(nf + 1) + 2 * (size_t)np)));
  int *ip = ints.as<int>();
  d.nw = ip;
  d.boff = ip + (nf + 1);
  d.loff = ip + 2 * (nf + 1);
  d.sfil = ip + 3 * (nf + 1);
  d.sseg = d.sfil + np;
  std::vector<int> hnw1(hnw);
  hnw1.push_back(0);
  FS_HIP(hipMemcpy(d.nw, hnw1.data(), sizeof(int) * (nf + 1), hipMemcpyHostToDevice));
  FS_HIP(hipMemcpy(d.boff, hboff.data(), sizeof(int) * (nf + 1), hipMemcpyHostToDevice));
  FS_HIP(hipMemcpy(d.loff, hloff.data(), sizeof(int) * (nf + 1), hipMemcpyHostToDevice));
  FS_HIP(hipMemcpy(d.sfil, sf.data(), sizeof(int) * np, hipMemcpyHostToDevice));
  FS_HIP(hipMemcpy(d.sseg, ss.data(), sizeof(int) * np, hipMemcpyHostToDevice));
  FS_TRY(joffb.alloc(sizeof(long) * (nf + 1)));
  FS_HIP(hipMemcpy(joffb.p, hjoff.data(), sizeof(long) * (nf + 1), hipMemcpyHostToDevice));
  d.joff = joffb.as<long>();
  FS_TRY(propsb.alloc(sizeof(fs_fil_props) * nf));
  FS_HIP(hipMemcpy(propsb.p, props, sizeof(fs_fil_props) * nf, hipMemcpyHostToDevice));
  d.props = propsb.as<fs_fil_props>();
  FS_TRY(Jb.alloc(sizeof(double) * (size_t)hjoff[nf]));
  max_nw = *std::max_element(hnw.begin(), hnw.end());
  if (banded()) {
    FS_TRY(Fb.alloc(sizeof(double) * ((size_t)np + 2 * (size_t)nf) * (kBlkRec / 2)));
  } else {
    FS_TRY(JTb.alloc(sizeof(double) * (size_t)hjoff[nf]));
    FS_TRY(Mb.alloc(sizeof(double) * (size_t)hjoff[nf]));
  }
  dot_chunk = 2048;
  dot_blocks = ceil_div((long)n6, dot_chunk);
  FS_TRY(alloc_history(p.broyden_maxiter));
  FS_TRY(flags.alloc(64));
  FS_TRY(ctl.alloc(4 * sizeof(int)));
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(4), 0, st, ctl.as<int>(), 4, 0);
  FS_LAUNCHED();
  FS_HIP(hipHostMalloc((void **)&hflags, 64, hipHostMallocDefault));
  d.np = np;
  d.nlam = nlam;
  d.nf = nf;
  d.dL = p.dL;
  d.dt = p.dt;
  d.mu = p.fcm.mu;
  d.Lx = p.fcm.nx * p.fcm.dx;
  d.Ly = p.fcm.ny * p.fcm.dx;
  d.Lz = p.fcm.nz * p.fcm.dx;
  d.impl_euler_steps = p.impl_euler_steps;
  d.sperm_decay = p.sperm_decay;
  d.st0 = p.strain_twist0;
  for (int k = 0; k < 3; ++k) d.weight[k] = p.weight_per_length[k];
  // FCM solver and collision barrier on the same box
  FS_TRY(fcm.init(p.fcm, np, p.precision));
  fs_colbar_params cp;
  cp.box[0] = d.Lx;
  cp.box[1] = d.Ly;
  cp.box[2] = d.Lz;
  cp.a = p.fcm.a;
  cp.fs_over_2a = p.barrier_fs_over_2a;
  cp.cap = p.barrier_cap;
  cp.initial_friction_steps = p.initial_friction_steps;
  if (p.barrier_on) FS_TRY(cb.init(cp, np, sf.data(), ss.data()));
  FS_HIP(hipStreamCreate(&st));
  return FS_OK;
}

void SystemImpl::release() {
  if (st) (void)hipStreamSynchronize(st);
  springb.release();
  arena.release();
  ints.release();
  joffb.release();
  propsb.release();
  Jb.release();
  Mb.release();
  JTb.release();
  Fb.release();
  Cm.release();
  Dm.release();
  partial.release();
  coefb.release();
  flags.release();
  ctl.release();
  if (hflags) (void)hipHostFree(hflags);
  hflags = nullptr;
  fcm.release();
  cb.release();
  if (st) (void)hipStreamDestroy(st);
  st = nullptr;
}

// ---------------------------------------------------------------------
// profiling: an event pair around each stage call, in the reference
// profiler's buckets; collected after the timestep's stream sync
// ---------------------------------------------------------------------
hipEvent_t SystemImpl::prof_event() {
  if (pev_used == pev.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    pev.push_back(e);
  }
  return pev[pev_used++];
}

void SystemImpl::prof_reset() {
  pev_used = 0;
  pmarks.clear();
  for (double &v : pms) v = 0;
  fcm.prof_used = 0;
}

int SystemImpl::prof_collect() {
  for (const PMark &m : pmarks) {
    float t = 0;
    if (m.a && m.z) FS_HIP(hipEventElapsedTime(&t, m.a, m.z));
    pms[m.b] += t;
  }
  double acc7[7] = {0, 0, 0, 0, 0, 0, 0};
  FS_TRY(fcm.prof_collect(acc7));
  pms[P_MAKEZ] += acc7[0];
  pms[P_GAUSS] += acc7[1];
  pms[P_FORCD] += acc7[2];
  pms[P_FFTFW] += acc7[3];
  pms[P_FLOWS] += acc7[4];
  pms[P_FFTBK] += acc7[5];
  pms[P_PARVR] += acc7[6];
  pmarks.clear();
  pev_used = 0;
  return FS_OK;
}

namespace {
struct ProfScope {
  SystemImpl &S;
  int b;
  hipEvent_t a = nullptr;
  ProfScope(SystemImpl &s, int bucket) : S(s), b(bucket) {
    if (S.prof) {
      a = S.prof_event();
      if (a) (void)hipEventRecord(a, S.st);
    }
  }
  ~ProfScope() {
    if (S.prof && a) {
      hipEvent_t z = S.prof_event();
      if (z) {
        (void)hipEventRecord(z, S.st);
        S.pmarks.push_back({b, a, z});
      }
    }
  }
};
}  // namespace

int SystemImpl::initial_guess() {
  ProfScope ps_(*this, P_ASSIGN);
  hipLaunchKernelGGL(k_guess, dim3(nf), dim3(64), 0, st, d);
  FS_LAUNCHED();
  return FS_OK;
}

// update(u) then forces(nt) in one launch
int SystemImpl::update_forces(const double *u, int nt) {
  ProfScope ps_(*this, P_ASSIGN);
  hipLaunchKernelGGL(k_update_forces, dim3(nf), dim3(64), 0, st, d, u, nt);
  FS_LAUNCHED();
  return springs();
}

int SystemImpl::forces(int nt) {
  ProfScope ps_(*this, P_ASSIGN);
  hipLaunchKernelGGL(k_forces, dim3(nf), dim3(64), 0, st, d, nt);
  FS_LAUNCHED();
  return springs();
}

// spring link forces (after the filament forces)
int SystemImpl::springs() {
  if (spring_nb > 0) {
    const int *sb = springb.as<int>();
    hipLaunchKernelGGL(k_springs, dim3(ceil_div(spring_nb, 64)), dim3(64), 0,
                       st, d, spring_nb, sb, sb + spring_nb,
                       sb + 2 * spring_nb + 1, sb + 2 * spring_nb + 1 + 2 * nspring,
                       spring_k, spring_len);
  }
  FS_LAUNCHED();
  return FS_OK;
}

// Spring links as (filament, bead, filament, bead) rows; per linked bead
// the list of its links in link order, uploaded as one int buffer:
// bead[nb] | start[nb + 1] | ent[2 n] | pair[2 n].
int SystemImpl::set_spring_links(int n, const int *links, double k,
                                 double len) {
  FS_CHECK(n >= 0 && (n == 0 || links), FS_ERR_ARG, "bad spring links");
  std::vector<int> pair(2 * (size_t)n);
  for (int l = 0; l < n; ++l)
    for (int e = 0; e < 2; ++e) {
      const int f = links[4 * l + 2 * e], b = links[4 * l + 2 * e + 1];
      FS_CHECK(f >= 0 && f < nf && b >= 0 && b < hnw[f], FS_ERR_ARG,
               "spring link %d: bead (%d, %d) is not in the system", l, f, b);
      pair[2 * l + e] = hboff[f] + b;
    }
  for (int l = 0; l < n; ++l)
    FS_CHECK(pair[2 * l] != pair[2 * l + 1], FS_ERR_ARG,
             "spring link %d joins a bead to itself", l);
  // entries sorted by bead, in link order within a bead
  std::vector<std::pair<int, int>> es;  // (bead, 2 link + second)
  es.reserve(2 * (size_t)n);
  for (int l = 0; l < n; ++l) {
    es.push_back({pair[2 * l], 2 * l});
    es.push_back({pair[2 * l + 1], 2 * l + 1});
  }
  std::stable_sort(es.begin(), es.end(),
                   [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                     return x.first < y.first;
                   });
  std::vector<int> bead, start, ent;
  for (size_t e = 0; e < es.size(); ++e) {
    if (e == 0 || es[e].first != es[e - 1].first) {
      bead.push_back(es[e].first);
      start.push_back((int)e);
    }
    ent.push_back(es[e].second);
  }
  start.push_back((int)es.size());
  const int nb = (int)bead.size();
  std::vector<int> buf;
  buf.insert(buf.end(), bead.begin(), bead.end());
  buf.insert(buf.end(), start.begin(), start.end());
  buf.insert(buf.end(), ent.begin(), ent.end());
  buf.insert(buf.end(), pair.begin(), pair.end());
  FS_HIP(hipStreamSynchronize(st));
  springb.release();
  spring_nb = 0;
  nspring = 0;
  if (n > 0) {
    FS_TRY(springb.alloc(sizeof(int) * buf.size()));
    FS_HIP(hipMemcpy(springb.p, buf.data(), sizeof(int) * buf.size(),
                     hipMemcpyHostToDevice));
  }
  spring_nb = nb;
  nspring = n;
  spring_k = k;
  spring_len = len;
  return FS_OK;
}

int SystemImpl::update(const double *u) {
  ProfScope ps_(*this, P_ASSIGN);
  hipLaunchKernelGGL(k_update, dim3(nf), dim3(64), 0, st, d, u);
  FS_LAUNCHED();
  return FS_OK;
}

// Jacobian blocks and their inverses; the singular flag stays on the device
int SystemImpl::build_jacobian_dev(int nt) {
  size_t nJ = (size_t)hjoff[nf];
  hipLaunchKernelGGL(k_fill_f64, dim3(std::min<size_t>(4096, (nJ + 255) / 256)),
                     dim3(256), 0, st, Jb.as<double>(), nJ, 0.0);
  hipLaunchKernelGGL(k_jac_build, dim3(nf), dim3(64), 0, st, d, nt,
                     Jb.as<double>());
  FS_LAUNCHED();
  int *sing = flags.as<int>() + 4;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, st, sing, 1, 0);
  if (banded()) {  // block-tridiagonal factorisation, one workgroup each
    hipLaunchKernelGGL(k_blk_factor, dim3(nf), dim3(kFacT), 0, st, d,
                       Jb.as<double>(), Fb.as<double>(), sing);
  } else {  // longer filaments: dense Gauss-Jordan in global memory (L2)
    FS_HIP(hipMemcpyAsync(Mb.p, Jb.p, sizeof(double) * nJ,
                          hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_gj_inverse, dim3(nf), dim3(1024), 0, st, d,
                       Mb.as<double>(), Jinv_cm(), sing);
  }
  FS_LAUNCHED();
  return FS_OK;
}

// Broyden history of `cols` columns (the step's own, plus carried ones):
// the dot partials and coefficients are sized for it, the 2 Broyden dot
// scalars follow the coefficients
int SystemImpl::alloc_history(int cols) {
  const size_t n6 = 6 * (size_t)np;
  Cm.release();
  Dm.release();
  partial.release();
  coefb.release();
  hcols = 0;
  FS_TRY(Cm.alloc(sizeof(double) * n6 * cols));
  FS_TRY(Dm.alloc(sizeof(double) * n6 * cols));
  FS_TRY(partial.alloc(sizeof(double) * (size_t)dot_blocks * std::max(2, cols)));
  FS_TRY(coefb.alloc(sizeof(double) * (cols + 2)));
  hcols = cols;
  return FS_OK;
}

int SystemImpl::set_history_carry(int cap) {
  FS_CHECK(cap >= 0 && cap <= kHistCarryMax, FS_ERR_ARG,
           "history carry must be in [0, %d]", kHistCarryMax);
  FS_HIP(hipStreamSynchronize(st));
  if (p.broyden_maxiter + cap > hcols) FS_TRY(alloc_history(p.broyden_maxiter + cap));
  hist_cap = cap;
  hist_cols = 0;
  return FS_OK;
}

int SystemImpl::build_jacobian(int nt, bool clear_history) {
  ProfScope ps_(*this, P_BLDJ);
  FS_TRY(build_jacobian_dev(nt));
  // history reset (FilamentJacobianSolver.hpp:40-41). The history kernels
  // read only columns below the count they are given, and a timestep's
  // Broyden iteration k reads columns 0 .. k-2, written by its own earlier
  // iterations: the timestep needs no clearing (neither does the batched
  // sweep's device-controlled step)
  if (clear_history) {
    size_t nh = 6 * (size_t)np * p.broyden_maxiter;
    hipLaunchKernelGGL(k_fill_f64, dim3(std::min<size_t>(4096, (nh + 255) / 256)),
                       dim3(256), 0, st, Cm.as<double>(), nh, 0.0);
    hipLaunchKernelGGL(k_fill_f64, dim3(std::min<size_t>(4096, (nh + 255) / 256)),
                       dim3(256), 0, st, Dm.as<double>(), nh, 0.0);
    FS_LAUNCHED();
  }
  int *sing = flags.as<int>() + 4;
  FS_HIP(hipMemcpyAsync((int *)hflags + 4, sing, sizeof(int),
                        hipMemcpyDeviceToHost, st));
  FS_HIP(hipStreamSynchronize(st));
  FS_CHECK(((int *)hflags)[4] == 0, FS_ERR_ARG,
           "approximate Jacobian is singular");
  return FS_OK;
}

// Dense J^-1 of filament f (row-major, host), the reference's inverse with
// its lambda rows scaled by 1/dt: the stored inverse, or with the banded
// factors one solve per unit vector (a test and inspection path).
int SystemImpl::jinv_dense(int f, double *out) {
  FS_HIP(hipStreamSynchronize(st));
  const int n = 6 * hnw[f];
  const size_t so = 6 * (size_t)hboff[f], n6 = 6 * (size_t)np;
  if (!banded()) {
    std::vector<double> m((size_t)n * n);
    FS_HIP(hipMemcpy(m.data(), JTb.as<double>() + hjoff[f],
                     sizeof(double) * m.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < n; ++r)  // stored column-major
      for (int c = 0; c < n; ++c) out[(size_t)r * n + c] = m[(size_t)c * n + r];
    return FS_OK;
  }
  dbuf work;
  FS_TRY(work.alloc(sizeof(double) * 2 * n6));
  double *b = work.as<double>(), *x = b + n6;
  std::vector<double> col(n);
  int rc = FS_OK;
  for (int c = 0; c < n && rc == FS_OK; ++c) {
    hipLaunchKernelGGL(k_fill_f64, dim3(std::min<size_t>(4096, (n6 + 255) / 256)),
                       dim3(256), 0, st, b, n6, 0.0);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(1), 0, st, b + so + c, (size_t)1,
                       1.0);
    rc = jsolve_impl(b, x, 0, nullptr, 0);
    if (rc == FS_OK &&
        hipMemcpyAsync(col.data(), x + so, sizeof(double) * n,
                       hipMemcpyDeviceToHost, st) != hipSuccess) {
      set_error("copy of an inverse column failed");
      rc = FS_ERR_HIP;
    }
    if (rc == FS_OK && hipStreamSynchronize(st) != hipSuccess) {
      set_error("inverse column solve failed");
      rc = FS_ERR_HIP;
    }
    for (int r = 0; r < n && rc == FS_OK; ++r) out[(size_t)r * n + c] = col[r];
  }
  work.release();
  return rc;
}

int SystemImpl::jsolve(const double *b, double *x, int iter) {
  FS_CHECK(iter >= 0 && iter < hcols, FS_ERR_ARG, "bad iter %d", iter);
  FS_CHECK(b != x, FS_ERR_ARG, "jsolve needs distinct b and x");
  return jsolve_impl(b, x, iter, nullptr, 0);
}

// x = blockdiag(J^-1) b + sum_{i < ncol} C_i (D_i . b). With ncolp the
// history length is *ncolp + noff, read on the device (grids sized for
// broyden_maxiter); the launches honour d.act.
int SystemImpl::jsolve_impl(const double *b, double *x, int ncol,
                            const int *ncolp, int noff) {
  ProfScope ps_(*this, P_SLVJ);
  const int *act = d.act;
  if (banded()) {
    hipLaunchKernelGGL(k_blk_solve, dim3(nf), dim3(64), 0, st, d,
                       Fb.as<double>(), b, x);
  } else {
    hipLaunchKernelGGL(k_jsolve_block, dim3(nf), dim3(256), 0, st, d,
                       Jinv_cm(), b, x);
  }
  FS_LAUNCHED();
  const int gy = ncolp ? p.broyden_maxiter : ncol;
  if (gy > 0) {
    size_t n6 = 6 * (size_t)np;
    hipLaunchKernelGGL(k_dots_partial, dim3(dot_blocks, gy), dim3(256), 0, st,
                       act, Dm.as<double>(), n6, ncol, ncolp, noff, gy, b, n6,
                       dot_chunk, partial.as<double>());
    hipLaunchKernelGGL(k_dots_final, dim3(gy), dim3(64), 0, st, act,
                       partial.as<double>(), dot_blocks, gy, ncol, ncolp, noff,
                       coefb.as<double>());
    hipLaunchKernelGGL(k_hist_axpy, dim3(ceil_div((long)n6, 256)), dim3(256), 0,
                       st, act, Cm.as<double>(), n6, ncol, ncolp, noff,
                       coefb.as<double>(), n6, x);
    FS_LAUNCHED();
  }
  return FS_OK;
}

int SystemImpl::broyden_update(const double *Eo, const double *En,
                               const double *upd, int iter) {
  FS_CHECK(iter >= 0 && iter < p.broyden_maxiter, FS_ERR_ARG, "bad iter %d",
           iter);
  size_t n6 = 6 * (size_t)np;
  hipLaunchKernelGGL(k_sub, dim3(ceil_div((long)n6, 256)), dim3(256), 0, st, En,
                     Eo, dE, n6);
  // ||dE||^2 through the same deterministic dot reduction
  hipLaunchKernelGGL(k_dots_partial, dim3(dot_blocks, 1), dim3(256), 0, st,
                     (const int *)nullptr, dE, n6, 1, (const int *)nullptr, 0,
                     1, dE, n6, dot_chunk, partial.as<double>());
  double *nrm2 = coefb.as<double>() + hcols;
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(64), 0, st,
                     (const int *)nullptr, partial.as<double>(), dot_blocks, 1,
                     1, (const int *)nullptr, 0, nrm2);
  FS_LAUNCHED();
  FS_TRY(jsolve(dE, tmp, iter));
  hipLaunchKernelGGL(k_broyden_cols, dim3(ceil_div((long)n6, 256)), dim3(256), 0,
                     st, tmp, upd, dE, nrm2, n6, Cm.as<double>() + n6 * iter,
                     Dm.as<double>() + n6 * iter);
  FS_LAUNCHED();
  return FS_OK;
}

int SystemImpl::residual(int nt, const double *V, const double *W, double *err,
                         int *check, double *maxabs) {
  ProfScope ps_(*this, P_ERRCHK);
  unsigned long long *mb = flags.as<unsigned long long>();
  hipLaunchKernelGGL(k_resid_reset, dim3(1), dim3(1), 0, st, d.act, mb);
  hipLaunchKernelGGL(k_residual, dim3(ceil_div(np, 256)), dim3(256), 0, st, d,
                     nt, V, W, err, mb);
  FS_LAUNCHED();
  FS_HIP(hipMemcpyAsync(hflags, mb, 4 * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, st));
  FS_HIP(hipStreamSynchronize(st));
  double m;
  memcpy(&m, hflags, sizeof m);
  if (maxabs) *maxabs = m;
  if (check) *check = (m > p.tol) ? 1 : 0;
  // MobilityConstraints.cpp:764-767 ("Prepare for explosion")
  if (m > 1e6) step_exploded = 1;
  if (hflags[3]) step_nonfinite = 1;
  return FS_OK;
}

int SystemImpl::step(int nt) {
  size_t n = std::max((size_t)np, 3 * (size_t)nlam);
  hipLaunchKernelGGL(k_step, dim3(ceil_div((long)n, 256)), dim3(256), 0, st, d,
                     nt);
  FS_LAUNCHED();
  return FS_OK;
}

int SystemImpl::mobility(int nt) {
  ProfScope ps_(*this, P_FCM);
  if (nt > p.initial_friction_steps)
    return fcm.mobility(np, d.Ya, d.Fa, d.Ta, d.V, d.W, st);
  return fcm.friction(np, d.Fa, d.Ta, d.V, d.W, st);
}

int SystemImpl::barrier(int nt, int link, double *min_dist2) {
  ProfScope ps_(*this, P_COLBAR);
  if (!p.barrier_on || barrier_off) {
    if (min_dist2) *min_dist2 = 100000.0;
    return FS_OK;
  }
  if (link) FS_TRY(cb.link(d.Xs, st));
  return cb.apply(d.Ya, nt, d.Fa, min_dist2, st);
}

// linked_list.F_total (CollisionBarrierFilament.cpp:681-686): the barrier
// force alone, which the step added to F. Recomputed into out (3 Np, device)
// from the current Y with the cell list linked at the step's start, each
// segment's pair sum added to zero as the step added it to the other forces;
// zero when the barrier is off.
int SystemImpl::barrier_force(int nt, double *out) {
  FS_CHECK(out || np == 0, FS_ERR_ARG, "null output");
  const size_t n3 = 3 * (size_t)np;
  if (n3 == 0) return FS_OK;
  hipLaunchKernelGGL(k_fill_f64, dim3(ceil_div((long)n3, 256)), dim3(256), 0,
                     st, out, n3, 0.0);
  FS_LAUNCHED();
  if (p.barrier_on && !barrier_off && cb.linked) {
    // a batch's activity flag may hold 0 for a converged system: apply
    // unconditionally here
    const int *act = cb.act;
    cb.act = nullptr;
    int rc = cb.apply(d.Ya, nt, out, nullptr, st);
    cb.act = act;
    FS_TRY(rc);
  }
  FS_HIP(hipStreamSynchronize(st));
  return FS_OK;
}

// main.cpp:284-754, with the two Jacobian solves of each Broyden iteration
// (main.cpp:518, 650) fused into one (k_broyden_fused).
int SystemImpl::timestep(int nt, int *iters) {
  FS_TRY(begin_step(nt));
  FS_TRY(broyden_loop(nt, iters));
  FS_TRY(step(nt));
  FS_HIP(hipStreamSynchronize(st));
  return FS_OK;
}

// main_gaitreset.cpp:310-812: the passes with gait_reset_i < 2 solve the
// step and skip Filament::step (:775-781), leaving the previous state.
int SystemImpl::timestep_trial(int nt, int *iters) {
  FS_TRY(begin_step(nt));
  FS_TRY(broyden_loop(nt, iters));
  FS_HIP(hipStreamSynchronize(st));
  return FS_OK;
}

int SystemImpl::set_barrier(int on, int mask) {
  FS_CHECK(on == 0 || on == 1, FS_ERR_ARG, "on must be 0 or 1");
  FS_CHECK(mask == FS_BARRIER_ALL || mask == FS_BARRIER_NOT_SWIMMER ||
               mask == FS_BARRIER_NOT_TEST_PARTICLES,
           FS_ERR_ARG, "unknown barrier mask %d", mask);
  FS_CHECK(on == 0 || p.barrier_on, FS_ERR_STATE,
           "the system was created with barrier_on = 0");
  if (p.barrier_on) cb.mask = mask;
  barrier_off = (on == 0);
  return FS_OK;
}

// Initial guess, forces, Jacobian, link + barrier, mobility, residual E.
int SystemImpl::begin_guess(int nt) {
  step_exploded = 0;
  step_nonfinite = 0;
  step_check = 1;
  FS_TRY(initial_guess());
  FS_TRY(forces(nt));
  FS_TRY(build_jacobian(nt, false));
  return barrier(nt, 1, nullptr);
}

int SystemImpl::eval_guess(int nt) {
  FS_TRY(mobility(nt));
  return residual(nt, d.V, d.W, E, &check0, &last_res0);
}

int SystemImpl::begin_step(int nt) {
  FS_TRY(begin_guess(nt));
  return eval_guess(nt);
}

int SystemImpl::broyden_loop(int nt, int *iters) {
  const size_t n6 = 6 * (size_t)np;
  int check = check0;
  int iter = 0;
  last_res = last_res0;
  // history columns carried in from the last step (0 without carrying)
  const int h0 = hist_cap > 0 ? hist_cols : 0;
  if (check) FS_TRY(jsolve(E, Upd, h0));
  while (check && iter < p.broyden_maxiter) {
    FS_TRY(update_forces(Upd, nt));
    FS_TRY(barrier(nt, 0, nullptr));
    FS_TRY(mobility(nt));
    FS_TRY(residual(nt, d.V, d.W, NewE, &check, &last_res));
    ++iter;
    if (!check || iter >= p.broyden_maxiter) break;
    // dE = NewE - E; |dE|^2 and dE.NewE; g = H_{iter-1} NewE
    double *dots = coefb.as<double>() + hcols;  // 2 scalars
    {
      ProfScope ps_(*this, P_SLVJ);
      hipLaunchKernelGGL(k_diff_dots, dim3(dot_blocks), dim3(256), 0, st,
                         (const int *)nullptr, NewE, E, dE, n6, dot_chunk,
                         partial.as<double>());
      hipLaunchKernelGGL(k_dots_final, dim3(2), dim3(64), 0, st,
                         (const int *)nullptr, partial.as<double>(), dot_blocks,
                         2, 2, (const int *)nullptr, 0, dots);
      FS_LAUNCHED();
    }
    FS_TRY(jsolve(NewE, tmp, h0 + iter - 1));
    {
      ProfScope ps_(*this, P_SLVJ);
      hipLaunchKernelGGL(k_broyden_fused, dim3(ceil_div((long)n6, 256)),
                         dim3(256), 0, st, (const int *)nullptr, tmp, dE, dots,
                         n6, Cm.as<double>(), Dm.as<double>(), h0 + iter - 1,
                         (const int *)nullptr, Upd);
      FS_LAUNCHED();
    }
    std::swap(E, NewE);
  }
  if (hist_cap > 0) {
    // columns now held: the carried ones and one per iteration that did
    // not end the loop. Beyond hist_cap the newest hist_cap / 2 move to the
    // front and the older ones are dropped
    const int held = h0 + std::max(0, iter - 1);
    if (held <= hist_cap) {
      hist_cols = held;
    } else {
      const int keep = hist_cap / 2;  // held > 2 keep: the ranges are disjoint
      const size_t cb = sizeof(double) * n6 * keep;
      if (keep > 0) {
        FS_HIP(hipMemcpyAsync(Cm.as<double>(), Cm.as<double>() + n6 * (held - keep), cb,
                              hipMemcpyDeviceToDevice, st));
        FS_HIP(hipMemcpyAsync(Dm.as<double>(), Dm.as<double>() + n6 * (held - keep), cb,
                              hipMemcpyDeviceToDevice, st));
      }
      hist_cols = keep;
    }
  }
  step_check = check;
  if (iters) *iters = iter;
  return FS_OK;
}

// ---------------------------------------------------------------------
// batched sweep: host-sync-free launch sequences of one timestep. Kernel
// arguments are fixed at capture; the step number and the loop state are
// read on the device (d.ntp, ctl), and every kernel of a Broyden iteration
// returns at once when the suspension has converged (d.act, fcm.act,
// cb.act). The arithmetic, launch order and reductions are those of
// timestep(): a batched suspension reproduces its unbatched run bit for bit.
// Requires initial_friction_steps == 0 (the FCM path on every step).
// ---------------------------------------------------------------------
void SystemImpl::set_device_control(const int *act, const int *ntp) {
  d.act = act;
  d.ntp = ntp;
  fcm.act = act;
  cb.act = act;
}

// part: -1 the whole sequence; 0 up to the mobility solve; 1 after it (a
// batched sweep with fused FCM stages solves every suspension's mobility
// between the two).
int SystemImpl::begin_step_dev(int *status, int part) {
  const int nt = 1;  // read from d.ntp on the device
  const int *ntp = d.ntp;
  set_device_control(nullptr, ntp);
  if (part != 1) {
    FS_TRY(initial_guess());
    FS_TRY(forces(nt));
    FS_TRY(build_jacobian_dev(nt));
    FS_TRY(barrier(nt, 1, nullptr));
  }
  if (part == -1) FS_TRY(mobility(nt));
  if (part == 0) return FS_OK;
  unsigned long long *mb = flags.as<unsigned long long>();
  hipLaunchKernelGGL(k_resid_reset, dim3(1), dim3(1), 0, st,
                     (const int *)nullptr, mb);
  hipLaunchKernelGGL(k_residual, dim3(ceil_div(np, 256)), dim3(256), 0, st, d,
                     nt, d.V, d.W, E, mb);
  hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(1), 0, st, ctl.as<int>(), mb,
                     p.tol, p.broyden_maxiter);
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, flags.as<int>() + 4, 1,
                     ctl.as<int>() + 2);
  FS_LAUNCHED();
  set_device_control(ctl.as<int>() + 1, ntp);
  FS_TRY(jsolve_impl(E, Upd, 0, nullptr, 0));
  set_device_control(nullptr, ntp);
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, ctl.as<int>(), 3,
                     status);
  FS_LAUNCHED();
  return FS_OK;
}

int SystemImpl::iteration_dev(int *status, int part) {
  const int nt = 1;  // read from d.ntp on the device
  const int *ntp = d.ntp;
  const int *act = ctl.as<int>() + 1;
  const size_t n6 = 6 * (size_t)np;
  set_device_control(act, ntp);
  if (part != 1) {
    FS_TRY(update_forces(Upd, nt));
    FS_TRY(barrier(nt, 0, nullptr));
  }
  if (part == -1) FS_TRY(mobility(nt));
  if (part == 0) {
    set_device_control(nullptr, ntp);
    return FS_OK;
  }
  unsigned long long *mb = flags.as<unsigned long long>();
  hipLaunchKernelGGL(k_resid_reset, dim3(1), dim3(1), 0, st, act, mb);
  hipLaunchKernelGGL(k_residual, dim3(ceil_div(np, 256)), dim3(256), 0, st, d,
                     nt, d.V, d.W, NewE, mb);
  hipLaunchKernelGGL(k_ctl_after, dim3(1), dim3(1), 0, st, ctl.as<int>(), mb,
                     p.tol, p.broyden_maxiter);
  double *dots = coefb.as<double>() + hcols;  // 2 scalars
  hipLaunchKernelGGL(k_diff_dots, dim3(dot_blocks), dim3(256), 0, st, act, NewE,
                     E, dE, n6, dot_chunk, partial.as<double>());
  hipLaunchKernelGGL(k_dots_final, dim3(2), dim3(64), 0, st, act,
                     partial.as<double>(), dot_blocks, 2, 2,
                     (const int *)nullptr, 0, dots);
  FS_LAUNCHED();
  FS_TRY(jsolve_impl(NewE, tmp, 0, ctl.as<int>(), -1));
  hipLaunchKernelGGL(k_broyden_fused, dim3(ceil_div((long)n6, 256)), dim3(256),
                     0, st, act, tmp, dE, dots, n6, Cm.as<double>(),
                     Dm.as<double>(), 0, ctl.as<int>(), Upd);
  hipLaunchKernelGGL(k_copy_f64, dim3(ceil_div((long)n6, 256)), dim3(256), 0, st,
                     act, NewE, E, n6);
  set_device_control(nullptr, ntp);
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, ctl.as<int>(), 3,
                     status);
  FS_LAUNCHED();
  return FS_OK;
}

int SystemImpl::end_step_dev(int *status) {
  (void)status;
  set_device_control(nullptr, d.ntp);
  return step(1);  // step number read from d.ntp
}

}  // namespace fs
