// warmstart.hpp -- learned first Broyden iterate attached to a system.
//
// No reference counterpart: replaces the extrapolated guess of
// Bead::initialGuess (Bead.hpp:202-217) by guess + MLP correction when that
// lowers the residual's infinity norm, else keeps the extrapolation.
#pragma once

#include "system_impl.hpp"

namespace fs {

struct MlpImpl;

// network sizes for Nw segments per filament (feature layout in
// warmstart.hip)
inline int ws_inputs(int nw) { return 24 * nw - 4; }
inline int ws_outputs(int nw) { return 6 * nw; }

struct WarmStart {
  MlpImpl *mlp = nullptr;
  bool recording = false;
  bool tracking = false;  // velocity/angular-velocity history kept (attached or recording)
  int nw = 0, nin = 0, nout = 0;
  dbuf V1, V2, W1, W2, feat, pred, u, sXs, sUs, sqs, slam;
  float *recX = nullptr, *recT = nullptr;
  int rec_cap = 0, rec_count = 0;
  int used_mlp = 0;
  double res_plain = 0, res_mlp = 0;  // res_plain NaN: not evaluated
  // periodic check (check_every > 1): the extrapolated guess is solved on
  // every check_every-th step and whenever the network's guess does not
  // beat ref_plain, the last extrapolation residual measured
  int check_every = 1, since_check = 0;
  double ref_plain = -1;

  bool attached() const { return tracking; }
  int setup(SystemImpl &S);  // sizes + buffers (uniform segment count)
  int attach(SystemImpl &S, MlpImpl *m);
  int record(SystemImpl &S, float *X, float *T, int capacity);
  int timestep(SystemImpl &S, int nt, int *iters);
  void release();
};

}  // namespace fs
