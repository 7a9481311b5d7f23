// warmstart.hip -- features, prediction and targets of the learned warm
// start, and the timestep that uses it (no reference counterpart; see
// warmstart.hpp).
//
// Per filament, everything is expressed in the frame of bead 0's current
// orientation q_0 (columns tangent, normal, binormal). Features (24 Nw - 4):
// per segment the tangent, the last converged velocity V1 and V1 - V2, the
// last converged angular velocity W1 and W1 - W2, and U - Ut (the change of
// the rotation increment over the last step); per link lam1 and
// lam1 - lam2; cos/sin of the driving phase at the new time. The target is (converged - extrapolated) for the unknowns
// [X_0, U_0..U_{Nw-1}, lam_0..lam_{Nw-2}], the Broyden state layout of
// Filament::update (Filament.hpp:425-452).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "warmstart.hpp"

#pragma clang fp contract(off)

namespace fs {

struct MlpImpl;
int mlp_forward(MlpImpl *m, int B, const float *X, float *Y, hipStream_t st);
int mlp_dims(MlpImpl *m, int *nin, int *nout);

__device__ __forceinline__ void frame_of(const double *q, double *t, double *n,
                                         double *b) {
  double q11 = 2 * q[1] * q[1], q22 = 2 * q[2] * q[2], q33 = 2 * q[3] * q[3];
  t[0] = 1 - q22 - q33;
  t[1] = 2 * (q[1] * q[2] + q[0] * q[3]);
  t[2] = 2 * (q[1] * q[3] - q[0] * q[2]);
  n[0] = 2 * (q[1] * q[2] - q[0] * q[3]);
  n[1] = 1 - q11 - q33;
  n[2] = 2 * (q[2] * q[3] + q[0] * q[1]);
  b[0] = 2 * (q[1] * q[3] + q[0] * q[2]);
  b[1] = 2 * (q[2] * q[3] - q[0] * q[1]);
  b[2] = 1 - q11 - q22;
}

struct Frame3 {
  double t[3], n[3], b[3];
  __device__ void to_local(const double *v, float *o) const {
    o[0] = (float)(t[0] * v[0] + t[1] * v[1] + t[2] * v[2]);
    o[1] = (float)(n[0] * v[0] + n[1] * v[1] + n[2] * v[2]);
    o[2] = (float)(b[0] * v[0] + b[1] * v[1] + b[2] * v[2]);
  }
  __device__ void to_world(const float *v, double *o) const {
    for (int d = 0; d < 3; ++d)
      o[d] = t[d] * (double)v[0] + n[d] * (double)v[1] + b[d] * (double)v[2];
  }
};

__device__ __forceinline__ Frame3 bead0_frame(const SysDev &S, int f) {
  Frame3 F;
  frame_of(S.q + 4 * S.boff[f], F.t, F.n, F.b);
  return F;
}

// Differences of nearly equal histories (V1 - V2, W1 - W2, U - Ut,
// lam1 - lam2) are formed in FP64 before the FP32 store: after
// normalisation and BF16 rounding the network could not recover them from
// the two terms.
__device__ __forceinline__ void local_diff(const Frame3 &F, const double *a,
                                           const double *b, float *o) {
  const double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  F.to_local(d, o);
}

__global__ void k_ws_features(SysDev S, int nt, int nw,
                              const double *__restrict__ V1,
                              const double *__restrict__ V2,
                              const double *__restrict__ W1,
                              const double *__restrict__ W2, int nin,
                              float *__restrict__ X) {
  const int f = blockIdx.x;
  const int b0 = S.boff[f], l0 = S.loff[f];
  const Frame3 F = bead0_frame(S, f);
  float *x = X + (size_t)f * nin;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) {
    const int g = b0 + i;
    double t[3], n[3], b[3];
    frame_of(S.q + 4 * g, t, n, b);
    F.to_local(t, x + 3 * i);
    F.to_local(V1 + 3 * g, x + 3 * nw + 3 * i);
    local_diff(F, V1 + 3 * g, V2 + 3 * g, x + 6 * nw + 3 * i);
    F.to_local(W1 + 3 * g, x + 9 * nw + 3 * i);
    local_diff(F, W1 + 3 * g, W2 + 3 * g, x + 12 * nw + 3 * i);
    local_diff(F, S.U + 3 * g, S.Ut + 3 * g, x + 15 * nw + 3 * i);
  }
  for (int l = threadIdx.x; l < nw - 1; l += blockDim.x) {
    const int g = l0 + l;
    F.to_local(S.lam1 + 3 * g, x + 18 * nw + 3 * l);
    local_diff(F, S.lam1 + 3 * g, S.lam2 + 3 * g,
               x + 18 * nw + 3 * (nw - 1) + 3 * l);
  }
  if (threadIdx.x == 0) {
    const fs_fil_props pr = S.props[f];
    double ph = pr.omega * (nt * S.dt) - pr.phase;
    x[nin - 2] = (float)cos(ph);
    x[nin - 1] = (float)sin(ph);
  }
}

// u = -(R delta): Filament::update subtracts u, so the state moves by +delta
__global__ void k_ws_apply(SysDev S, int nw, const float *__restrict__ P,
                           int nout, double *__restrict__ u) {
  const int f = blockIdx.x;
  const Frame3 F = bead0_frame(S, f);
  const size_t so = 6 * (size_t)S.boff[f];
  for (int k = threadIdx.x; k < 2 * nw; k += blockDim.x) {
    double w[3];
    F.to_world(P + (size_t)f * nout + 3 * k, w);
    for (int d = 0; d < 3; ++d) u[so + 3 * k + d] = -w[d];
  }
}

// target = converged - guess in the local frame, same layout as the update
__global__ void k_ws_target(SysDev S, int nw, const double *__restrict__ gXs,
                            const double *__restrict__ gUs,
                            const double *__restrict__ glam, int nout,
                            float *__restrict__ T) {
  const int f = blockIdx.x;
  const int b0 = S.boff[f], l0 = S.loff[f];
  const Frame3 F = bead0_frame(S, f);
  float *t = T + (size_t)f * nout;
  for (int k = threadIdx.x; k < 2 * nw; k += blockDim.x) {
    double v[3];
    if (k == 0) {
      for (int d = 0; d < 3; ++d) v[d] = S.Xs[3 * b0 + d] - gXs[3 * b0 + d];
    } else if (k <= nw) {
      int g = b0 + k - 1;
      for (int d = 0; d < 3; ++d) v[d] = S.Us[3 * g + d] - gUs[3 * g + d];
    } else {
      int l = l0 + k - nw - 1;
      for (int d = 0; d < 3; ++d) v[d] = S.lam[3 * l + d] - glam[3 * l + d];
    }
    F.to_local(v, t + 3 * k);
  }
}

int WarmStart::setup(SystemImpl &S) {
  if (nw) return FS_OK;
  for (int f = 1; f < S.nf; ++f)
    FS_CHECK(S.hnw[f] == S.hnw[0], FS_ERR_ARG,
             "the warm start needs one segment count for every filament");
  nw = S.hnw[0];
  nin = ws_inputs(nw);
  nout = ws_outputs(nw);
  const size_t n3 = 3 * (size_t)S.np, n4 = 4 * (size_t)S.np,
               l3 = 3 * (size_t)S.nlam;
  FS_TRY(V1.alloc(sizeof(double) * n3));
  FS_TRY(V2.alloc(sizeof(double) * n3));
  FS_TRY(W1.alloc(sizeof(double) * n3));
  FS_TRY(W2.alloc(sizeof(double) * n3));
  FS_HIP(hipMemset(V1.p, 0, sizeof(double) * n3));
  FS_HIP(hipMemset(V2.p, 0, sizeof(double) * n3));
  FS_HIP(hipMemset(W1.p, 0, sizeof(double) * n3));
  FS_HIP(hipMemset(W2.p, 0, sizeof(double) * n3));
  FS_TRY(feat.alloc(sizeof(float) * (size_t)S.nf * nin));
  FS_TRY(pred.alloc(sizeof(float) * (size_t)S.nf * nout));
  FS_TRY(u.alloc(sizeof(double) * 6 * (size_t)S.np));
  FS_HIP(hipMemset(u.p, 0, sizeof(double) * 6 * (size_t)S.np));
  FS_TRY(sXs.alloc(sizeof(double) * n3));
  FS_TRY(sUs.alloc(sizeof(double) * n3));
  FS_TRY(sqs.alloc(sizeof(double) * n4));
  FS_TRY(slam.alloc(sizeof(double) * std::max<size_t>(l3, 1)));
  tracking = true;
  return FS_OK;
}

int WarmStart::attach(SystemImpl &S, MlpImpl *m) {
  if (!m) {
    mlp = nullptr;
    tracking = recording;
    return FS_OK;
  }
  FS_TRY(setup(S));
  int a = 0, b = 0;
  FS_TRY(mlp_dims(m, &a, &b));
  FS_CHECK(a == nin && b == nout, FS_ERR_ARG,
           "MLP is %d -> %d, the system needs %d -> %d", a, b, nin, nout);
  mlp = m;
  tracking = true;
  since_check = 0;
  ref_plain = -1;  // the first step solves both guesses
  return FS_OK;
}

int WarmStart::record(SystemImpl &S, float *X, float *T, int capacity) {
  if (!X || !T || capacity <= 0) {
    recording = false;
    recX = recT = nullptr;
    tracking = mlp != nullptr;
    return FS_OK;
  }
  FS_TRY(setup(S));
  recX = X;
  recT = T;
  rec_cap = capacity;
  rec_count = 0;
  recording = true;
  tracking = true;
  return FS_OK;
}

void WarmStart::release() {
  V1.release();
  V2.release();
  W1.release();
  W2.release();
  feat.release();
  pred.release();
  u.release();
  sXs.release();
  sUs.release();
  sqs.release();
  slam.release();
}

int WarmStart::timestep(SystemImpl &S, int nt, int *iters) {
  hipStream_t st = S.st;
  const size_t n3 = 3 * (size_t)S.np, n4 = 4 * (size_t)S.np,
               l3 = 3 * (size_t)S.nlam, n6 = 6 * (size_t)S.np;
  // the extrapolated guess is solved on check steps (every one by default)
  // and, on the others, only when the network's guess is rejected
  const bool check = !mlp || check_every <= 1 || ref_plain < 0 ||
                     since_check + 1 >= check_every;
  if (check) {
    FS_TRY(S.begin_step(nt));
    res_plain = S.last_res0;
  } else {
    FS_TRY(S.begin_guess(nt));
    res_plain = std::nan("");
  }
  res_mlp = 0;
  used_mlp = 0;
  hipLaunchKernelGGL(k_ws_features, dim3(S.nf), dim3(64), 0, st, S.d, nt,
                     nw, V1.as<double>(), V2.as<double>(), W1.as<double>(),
                     W2.as<double>(), nin, feat.as<float>());
  FS_LAUNCHED();
  // keep the extrapolated guess: fallback state and training reference
  FS_HIP(hipMemcpyAsync(sXs.p, S.d.Xs, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
  FS_HIP(hipMemcpyAsync(sUs.p, S.d.Us, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
  FS_HIP(hipMemcpyAsync(sqs.p, S.d.qs, sizeof(double) * n4, hipMemcpyDeviceToDevice, st));
  if (l3)
    FS_HIP(hipMemcpyAsync(slam.p, S.d.lam, sizeof(double) * l3, hipMemcpyDeviceToDevice, st));
  if (mlp && (!check || S.check0)) {
    FS_TRY(mlp_forward(mlp, S.nf, feat.as<float>(), pred.as<float>(), st));
    hipLaunchKernelGGL(k_ws_apply, dim3(S.nf), dim3(64), 0, st, S.d, nw,
                       pred.as<float>(), nout, u.as<double>());
    FS_LAUNCHED();
    FS_TRY(S.update_forces(u.as<double>(), nt));
    FS_TRY(S.barrier(nt, 0, nullptr));
    FS_TRY(S.mobility(nt));
    int chk = 0;
    FS_TRY(S.residual(nt, S.d.V, S.d.W, S.NewE, &chk, &res_mlp));
    if (check ? res_mlp < res_plain : res_mlp < ref_plain) {
      std::swap(S.E, S.NewE);
      S.check0 = chk;
      S.last_res0 = res_mlp;
      used_mlp = 1;
    } else {  // fall back to the extrapolated guess
      FS_HIP(hipMemcpyAsync(S.d.Xs, sXs.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
      FS_HIP(hipMemcpyAsync(S.d.Us, sUs.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
      FS_HIP(hipMemcpyAsync(S.d.qs, sqs.p, sizeof(double) * n4, hipMemcpyDeviceToDevice, st));
      if (l3)
        FS_HIP(hipMemcpyAsync(S.d.lam, slam.p, sizeof(double) * l3, hipMemcpyDeviceToDevice, st));
      if (!check) {
        // not solved yet: the forces, barrier, mobility and residual of the
        // restored guess, as begin_step computes them (same state, same
        // cell list)
        FS_TRY(S.forces(nt));
        FS_TRY(S.barrier(nt, 0, nullptr));
        FS_TRY(S.eval_guess(nt));
        res_plain = S.last_res0;
      }
    }
  }
  if (std::isnan(res_plain)) {
    ++since_check;
  } else {
    ref_plain = res_plain;
    since_check = 0;
  }
  FS_TRY(S.broyden_loop(nt, iters));
  if (recording && rec_count + S.nf <= rec_cap) {
    hipLaunchKernelGGL(k_ws_target, dim3(S.nf), dim3(64), 0, st, S.d, nw,
                       sXs.as<double>(), sUs.as<double>(), slam.as<double>(),
                       nout, recT + (size_t)rec_count * nout);
    FS_HIP(hipMemcpyAsync(recX + (size_t)rec_count * nin, feat.p,
                          sizeof(float) * (size_t)S.nf * nin,
                          hipMemcpyDeviceToDevice, st));
    FS_LAUNCHED();
    rec_count += S.nf;
  }
  // velocity and angular-velocity history for the next step's features
  FS_HIP(hipMemcpyAsync(V2.p, V1.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
  FS_HIP(hipMemcpyAsync(V1.p, S.d.V, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
  FS_HIP(hipMemcpyAsync(W2.p, W1.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
  FS_HIP(hipMemcpyAsync(W1.p, S.d.W, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
  (void)n6;
  FS_TRY(S.step(nt));
  FS_HIP(hipStreamSynchronize(st));
  return FS_OK;
}

}  // namespace fs
