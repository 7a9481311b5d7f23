// fcm_impl.hpp -- host-side state of one FCM solver (struct fs_fcm).
#pragma once

#include <vector>

#include "common.hpp"
#include "fft.hpp"

namespace fs {

// Geometry of the gather spread: grid tiles of tx*ty*tz points and the
// bins (same tiles) that hold the stencil origins able to reach a tile.
struct SpreadGeom {
  int nx, ny, nz, ngd;
  int tx, ty, tz;     // tile edge
  int nbx, nby, nbz;  // tiles per dimension
  int hx, hy, hz;     // neighbouring bins per dimension (besides own)
};

struct FcmImpl {
  fs_fcm_params p{};
  int max_np = 0;
  int precision = FS_FP64;
  int tsz = 8;  // bytes per real
  int last_np = 0;
  int timing = 0;
  // accumulate stage times over every solve (profiling): one set of 7
  // events per solve from a pool, summed by prof_collect
  int prof = 0;
  std::vector<hipEvent_t> prof_ev;
  size_t prof_used = 0;
  int prof_collect(double *acc7);
  FftPlan plx{}, ply{}, plz{};
  fs_fcm_tuning tune{};
  // spread tile geometry (mirrors SpreadGeom in fcm.hip)
  SpreadGeom geom{};
  int nbins = 0;
  bool mfma_spread = false;  // spread on the matrix cores (8 x 8 x 16 tiles)
  int grid_interp = 0;  // blocks of the persistent interpolation
  int zP = 1, yB = 1, xQ = 1;  // FFT pass workgroup shapes
  // device flag read by every kernel of a solve: 0 = skip (a converged
  // suspension of a batched sweep); nullptr = always run
  const int *act = nullptr;
  dbuf R, C, twx, twy, twz;
  dbuf tab, origin, coef, bin, slot, ids, bin_count, bin_start, scan_tmp;
  // matrix-core path: segment records and packed origins in bin order,
  // per-tile record lists (slab of kTileCap per tile, count, pool block of
  // the rest, the pool) and the interpolation's partial sums
  // [segment][component][tile slot]
  dbuf srec, sorg, tslab, tcnt, tovf, tpool, ipart;
  hipEvent_t ev[8] = {};
  hipEvent_t kev[11] = {};  // kernel boundaries of a timed solve
  bool kev_used[11] = {};

  int init(const fs_fcm_params &pp, int maxnp, int prec);
  int set_geometry();
  void release();
  int mobility(int np, const double *Y, const double *F, const double *Tq,
               double *V, double *W, hipStream_t st);
  int friction(int np, const double *F, const double *Tq, double *V,
               double *W, hipStream_t st);
  int stage_times(float *ms7);
  int kernel_times(float *ms10);
  int set_tuning(const fs_fcm_tuning &t);
  int autotune(int np, int reps, fs_fcm_tuning *best);
  template <typename T>
  int run(int np, const double *Y, const double *F, const double *Tq,
          double *V, double *W, hipStream_t st);
};

// Mobility solves of several suspensions with one launch per FCM stage
// (batched sweep). Members need equal launch shapes (fcm_same_shapes) and
// the matrix-core tile path; build() fills the per-member argument tables
// in device memory once (fixed Y/F/T/V/W and loop-control pointers), then
// each solve() launches every stage once over the whole group.
struct FcmGroupMember {
  FcmImpl *fcm = nullptr;
  const int *act = nullptr;  // loop-control flag (nullptr: always run)
  int np = 0;
  const double *Y = nullptr, *F = nullptr, *T = nullptr;
  double *V = nullptr, *W = nullptr;
};
struct FcmGroup {
  std::vector<FcmGroupMember> m;
  dbuf tab;                 // argument tables
  std::vector<size_t> off;  // offset of each stage's table
  int build();
  int solve(hipStream_t st);
  void release() { tab.release(); }
};
bool fcm_same_shapes(const FcmImpl &a, const FcmImpl &b);

}  // namespace fs
