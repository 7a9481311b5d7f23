// system_impl.hpp -- device state of a filament system (fs_system).
#pragma once

#include <vector>

#include "colbar_impl.hpp"
#include "common.hpp"
#include "fcm_impl.hpp"

namespace fs {

constexpr int kHistCarryMax = 192;  // Broyden columns a timestep may carry in

// Plain struct of device pointers and scalars passed by value to kernels.
struct SysDev {
  double *X, *Xs, *Xt, *U, *Us, *Ut, *q, *qs, *qt, *F, *TAU;
  double *lam, *lam1, *lam2;
  double *st;              // myStrainTwist per filament (3)
  double *Ya, *Fa, *Ta;    // FCM inputs (assign_filament_data)
  double *V, *W;           // FCM outputs
  int *nw, *boff, *loff, *sfil, *sseg;
  long *joff;
  fs_fil_props *props;
  int np, nlam, nf;
  double dL, dt, mu, Lx, Ly, Lz, st0;
  double weight[3];
  int impl_euler_steps, sperm_decay;
  // batched sweep only (nullptr otherwise): kernels of a Broyden iteration
  // return at once while *act == 0; *ntp replaces the step number argument
  const int *act, *ntp;
};

struct SystemImpl {
  fs_sim_params p{};
  int nf = 0, np = 0, nlam = 0;
  std::vector<int> hnw, hboff, hloff;
  std::vector<long> hjoff;
  SysDev d{};
  dbuf arena, ints, joffb, propsb, Jb, Mb, JTb, Cm, Dm, partial, coefb, flags;
  // block-tridiagonal factors of the Jacobians, used while every filament
  // has at most 64 segments (the solve keeps a filament's factors in LDS);
  // the dense inverse JTb (Gauss-Jordan) otherwise
  dbuf Fb;
  bool banded() const { return max_nw <= 64; }
  double *E = nullptr, *NewE = nullptr, *Upd = nullptr, *tmp = nullptr,
         *dE = nullptr, *scratch6 = nullptr;
  unsigned long long *hflags = nullptr;
  int dot_chunk = 0, dot_blocks = 0;
  int max_nw = 0;
  double last_res0 = 0, last_res = 0;
  // Broyden history carried into the next timestep (no reference
  // counterpart; the reference restarts it each step): up to hist_cap
  // columns, hist_cols of them carried into the current step; 0 = restart
  int hist_cap = 0, hist_cols = 0;
  int hcols = 0;  // history column capacity (broyden_maxiter + the carry cap)
  int alloc_history(int cols);
  int set_history_carry(int cap);
  FcmImpl fcm;
  ColbarImpl cb;
  hipStream_t st = nullptr;

  double *Jinv_cm() { return JTb.as<double>(); }
  int init(const fs_sim_params &pp, int nfil, const int *nw,
           const fs_fil_props *props);
  void release();
  int initial_guess();
  int forces(int nt);
  int update(const double *u);
  int update_forces(const double *u, int nt);  // update + forces, one launch
  int springs();
  // clear_history: zero the Broyden history (the API's reset); the
  // timestep reads only the columns its own iterations wrote and skips it
  int build_jacobian(int nt, bool clear_history = true);
  int jsolve(const double *b, double *x, int iter);
  int jinv_dense(int f, double *out);
  int broyden_update(const double *Eo, const double *En, const double *upd,
                     int iter);
  int residual(int nt, const double *V, const double *W, double *err,
               int *check, double *maxabs);
  int step(int nt);
  int mobility(int nt);
  int barrier(int nt, int link, double *min_dist2);
  int barrier_force(int nt, double *out);
  // barrier of the gait-reset passes: skipped while barrier_off is set
  bool barrier_off = false;
  int set_barrier(int on, int mask);
  int timestep(int nt, int *iters);
  int timestep_trial(int nt, int *iters);
  int begin_step(int nt);  // begin_guess + eval_guess
  // the step's extrapolated guess, forces, Jacobian and cell list
  int begin_guess(int nt);
  // mobility and residual of the current guess into E, check0, last_res0
  int eval_guess(int nt);
  int broyden_loop(int nt, int *iters);
  int check0 = 0;
  // outcome of the last step: the final residual still above TOL, a
  // residual above 1e6, a residual holding NaN or infinity
  int step_check = 0, step_exploded = 0, step_nonfinite = 0;

  // spring links between beads (spring_link_functions.cpp:24-45)
  dbuf springb;
  int nspring = 0, spring_nb = 0;
  double spring_k = 0, spring_len = 0;
  int set_spring_links(int n, const int *links, double k, double len);

  // --- profiling: the reference profiler's buckets (profilers.hpp:24-57),
  // accumulated per timestep from HIP event pairs on the system's stream
  enum {
    P_FCM, P_ASSIGN, P_BLDJ, P_ERRCHK, P_SLVJ, P_TOTAL, P_COLBAR,
    P_MAKEZ, P_GAUSS, P_FORCD, P_FFTFW, P_FFTBK, P_FLOWS, P_PARVR, P_N
  };
  int prof = 0;
  std::vector<hipEvent_t> pev;
  size_t pev_used = 0;
  struct PMark {
    int b;
    hipEvent_t a, z;
  };
  std::vector<PMark> pmarks;
  double pms[P_N] = {};
  hipEvent_t prof_event();
  void prof_reset();
  int prof_collect();

  // --- batched sweep: the timestep as three host-sync-free launch
  // sequences, captured once into graphs and replayed. ctl = {iterations,
  // still iterating, singular Jacobian}; ntp = device step number.
  dbuf ctl;
  int jsolve_impl(const double *b, double *x, int ncol, const int *ncolp,
                  int noff);
  int build_jacobian_dev(int nt);
  void set_device_control(const int *act, const int *ntp);
  int begin_step_dev(int *status, int part = -1);  // status: where to publish ctl
  int iteration_dev(int *status, int part = -1);
  int end_step_dev(int *status);
};

}  // namespace fs
